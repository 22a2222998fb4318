"""GPU tests of the batched many-fit path (kmd_kmeans_lloyd_batched): every
fit of a batch must give exactly what kmeans_lloyd_driver gives it alone --
labels, centroids (bit for bit), iteration count and inertia -- whatever the
mix of shapes, convergence times, empty-cluster relocations and the number of
graph replays between host checks."""
import numpy as np
import pytest
from numpy.testing import assert_array_equal

from oracle import lloyd_oracle as lo
from tests.conftest import to_dev, to_np

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def make_fit(rng, n, d, k, seed, dup=False):
    Xt, init = lo.gaussian_blobs(n, d, k, seed=seed, std=float(rng.uniform(0.5, 2.0)))
    if dup:  # two identical initial centroids: one cluster starts empty
        init = init.copy()
        init[:, 1] = init[:, 0]
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    return Xt, w, init


def single(kd, Xt, w, init, tol, max_iter):
    dev = torch.device("cuda", 0)
    init_d = to_dev(init)
    res_d = torch.empty_like(init_d)
    lab_d = torch.empty(Xt.shape[1], dtype=torch.int32, device=dev)
    n_iter, inertia, _ = kd.kmeans_lloyd_driver_ex(
        to_dev(Xt), to_dev(w), init_d, lab_d, res_d, tol, False, max_iter, "fp32",
        torch.cuda.current_stream())
    return to_np(lab_d), to_np(res_d), int(n_iter), float(inertia)


def batched(kd, problems, replays, precision="fp32"):
    dev = torch.device("cuda", 0)
    fits, outs = [], []
    for Xt, w, init, tol, max_iter in problems:
        init_d = to_dev(init)
        res_d = torch.empty_like(init_d)
        lab_d = torch.empty(Xt.shape[1], dtype=torch.int32, device=dev)
        fits.append((to_dev(Xt), to_dev(w), init_d, lab_d, res_d, tol, max_iter))
        outs.append((lab_d, res_d))
    results, stats = kd.kmeans_lloyd_batched(fits, precision, torch.cuda.current_stream(),
                                             replays)
    return [(to_np(l), to_np(r), int(it), float(inr))
            for (l, r), (it, inr) in zip(outs, results)], stats


def problem_mix(seed):
    rng = np.random.default_rng(seed)
    shapes = [(1000, 16, 8), (4097, 32, 32), (20000, 64, 128), (30001, 16, 256),
              (2500, 5, 3), (12000, 40, 64), (7000, 64, 256), (5000, 32, 100),
              (9000, 16, 32), (15000, 48, 8), (3000, 64, 64), (6000, 7, 33)]
    probs = []
    for i, (n, d, k) in enumerate(shapes):
        Xt, w, init = make_fit(rng, n, d, k, seed=100 * seed + i, dup=(i % 4 == 1))
        max_iter = [40, 1, 0, 60][i % 4] if i < 4 else 50
        probs.append((Xt, w, init, 1e-4, max_iter))
    return probs


@pytest.mark.parametrize("replays,precision", [(1, "fp32"), (4, "fp32"), (4, "bf16")])
def test_batched_fits_equal_single_fits(kd, torch_dev, replays, precision):
    """FP32 and BF16 (+ certified FP32 re-score) batched assignment: every fit
    bitwise equal to the FP32 single-fit driver."""
    probs = problem_mix(seed=1)
    refs = [single(kd, Xt, w, init, tol, mi) for Xt, w, init, tol, mi in probs]
    got, stats = batched(kd, probs, replays, precision)
    assert stats["graph_launches"] >= max(r[2] for r in refs)
    for i, (g, r) in enumerate(zip(got, refs)):
        assert g[2] == r[2], (i, g[2], r[2])
        assert_array_equal(g[0], r[0])
        assert_array_equal(g[1].view(np.uint32), r[1].view(np.uint32))
        assert g[3] == r[3], (i, g[3], r[3])


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_batched_relocates_like_the_driver(kd, torch_dev, precision):
    """Duplicate initial centroids leave clusters empty after the first step;
    the batched finalize relocates them to the same farthest samples."""
    rng = np.random.default_rng(9)
    probs = []
    for i in range(6):
        Xt, w, init = make_fit(rng, 5000 + 37 * i, 16 + 8 * i, 16, seed=500 + i)
        init = init.copy()
        if i % 2 == 0:
            init[:, 1:4] = init[:, [0]]  # three duplicates of centroid 0
        else:
            init[:, :] = init[:, [0]]    # all equal: k - 1 empty clusters
        probs.append((Xt, w, init, 0.0, 5))
    refs = [single(kd, Xt, w, init, tol, mi) for Xt, w, init, tol, mi in probs]
    got, _ = batched(kd, probs, 2, precision)
    for g, r in zip(got, refs):
        assert g[2] == r[2]
        assert_array_equal(g[0], r[0])
        assert_array_equal(g[1].view(np.uint32), r[1].view(np.uint32))
        assert g[3] == r[3]


def test_batched_mixes_in_unserved_shapes(kd, torch_dev):
    """A fit beyond the batched kernels' shapes (d = 100) goes through the
    single-fit driver inside the same call."""
    rng = np.random.default_rng(3)
    probs = [make_fit(rng, 4000, 100, 20, seed=77) + (1e-4, 20),
             make_fit(rng, 3000, 16, 10, seed=78) + (1e-4, 20)]
    refs = [single(kd, Xt, w, init, tol, mi) for Xt, w, init, tol, mi in probs]
    got, _ = batched(kd, probs, 4)
    for g, r in zip(got, refs):
        assert g[2] == r[2]
        assert_array_equal(g[0], r[0])
        assert_array_equal(g[1].view(np.uint32), r[1].view(np.uint32))


@pytest.mark.parametrize("replays,precision", [(1, "fp32"), (3, "fp32"), (3, "bf16")])
def test_batched_edge_shapes_equal_single_fits(kd, torch_dev, replays, precision):
    """Shapes at the segment sums' and assignment's edges: d = 1 (128 row
    groups of two columns), d = 63 (four groups of 64), d = 64 (three groups
    of 65), k = 1 (one cluster split over many 256-member items), n below one
    32-sample column block, n = k, and a ragged last 4096-sample tile."""
    rng = np.random.default_rng(21)
    shapes = [(40000, 1, 2), (3000, 63, 5), (8191, 64, 17), (20000, 3, 1),
              (20, 3, 4), (64, 2, 64), (4097, 16, 256), (1, 1, 1)]
    probs = []
    for i, (n, d, k) in enumerate(shapes):
        Xt, w, init = make_fit(rng, n, d, k, seed=900 + i)
        probs.append((Xt, w, init, 1e-4, 30))
    refs = [single(kd, Xt, w, init, tol, mi) for Xt, w, init, tol, mi in probs]
    got, _ = batched(kd, probs, replays, precision)
    for i, (g, r) in enumerate(zip(got, refs)):
        assert g[2] == r[2], (shapes[i], g[2], r[2])
        assert_array_equal(g[0], r[0])
        assert_array_equal(g[1].view(np.uint32), r[1].view(np.uint32))
        assert g[3] == r[3], (shapes[i], g[3], r[3])


def test_wide_range_fit_in_a_batch_goes_to_the_single_fit_driver(kd, torch_dev):
    """A fit whose data range needs floating-point sums (one outlier of 1e6
    beside a cluster of values below 1e-3) is taken out of the batched
    kernels and fitted by the single-fit driver; the other fits stay
    bitwise equal to their single-fit runs."""
    from tests.test_gpu_driver import _wide_range_data
    Xw, initw, groups = _wide_range_data(seed=23)
    probs = problem_mix(seed=4)[:4]
    probs.insert(2, (Xw, np.ones(Xw.shape[1], np.float32), initw, 0.0, 8))
    refs = [single(kd, Xt, w, init, tol, mi) for Xt, w, init, tol, mi in probs]
    got, _ = batched(kd, probs, 4)
    for i, (g, r) in enumerate(zip(got, refs)):
        assert g[2] == r[2], i
        assert_array_equal(g[0], r[0])
        if i != 2:
            assert_array_equal(g[1].view(np.uint32), r[1].view(np.uint32))
    lab, C = got[2][0], got[2][1]
    assert_array_equal(lab, groups)
    for c in range(6):
        mean = Xw[:, lab == c].astype(np.float64).mean(axis=1)
        assert np.abs(C[:, c] - mean).max() <= 1e-5 * np.abs(mean).max()
    # the continuous-batching engine refuses it with a message that the
    # request loop uses to route it to the single-fit driver
    eng = kd.Engine()
    dev = torch.device("cuda", 0)
    with pytest.raises(ValueError, match="single-fit driver"):
        eng.add(to_dev(Xw), to_dev(np.ones(Xw.shape[1], np.float32)), to_dev(initw),
                torch.empty(Xw.shape[1], dtype=torch.int32, device=dev),
                torch.empty(initw.shape, device=dev), 0.0, 8)
    del eng


def test_batched_bf16_with_near_ties_equals_fp32(kd, torch_dev):
    """Duplicate and near-duplicate initial centroids make every sample of
    their blobs a near-tie: the BF16 pass leaves those to the FP32 re-score,
    and the fits stay bitwise equal to the FP32 single-fit driver."""
    rng = np.random.default_rng(13)
    probs = []
    for i, (n, d, k) in enumerate([(6000, 16, 32), (9000, 32, 64), (7000, 64, 128)]):
        Xt, w, init = make_fit(rng, n, d, k, seed=1300 + i)
        init = init.copy()
        init[:, 1::2] = init[:, 0::2] + np.float32(1e-4)   # pairs 1e-4 apart
        probs.append((Xt, w, init, 1e-4, 25))
    refs = [single(kd, Xt, w, init, tol, mi) for Xt, w, init, tol, mi in probs]
    got, _ = batched(kd, probs, 4, "bf16")
    for g, r in zip(got, refs):
        assert g[2] == r[2]
        assert_array_equal(g[0], r[0])
        assert_array_equal(g[1].view(np.uint32), r[1].view(np.uint32))
        assert g[3] == r[3]
