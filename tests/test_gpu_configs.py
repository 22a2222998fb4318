"""GPU parity tests at the full sizes of the benchmark workloads, against the
oracle (the C restatement of the reference's assignment, FP64 NumPy updates
and inertia over the GPU's own labels) and against the single-fit driver:

* plain Lloyd, n = 10,000,000, d = 32, k = 256 (blobs: centres in
  [-20, 20], per-blob std in [0.5, 2]), and the same data shape with k = 8
  and 32;
* n = 2,000,000, d = 128, k = 1024 anisotropic blobs, FP32 and the BF16
  cross term with exact re-score;
* 512 independent fits in one batched call;
* a served stream with large requests of n = 2,000,000, d = 96, k = 512.

Each step of a fit is pinned separately so that rounding can never send
the GPU and the oracle down different trajectories: the labels a driver
returns are those of its final centroids (restatement, near-ties counted
and bounded), the centroids after i + 1 iterations are the FP64 means of
the labels of the centroids after i iterations (<= 1e-5 relative, the
reference driver's update, src/kmeans_lloyd_driver.hpp:83-354), counts are
exact integers, and the inertia is the FP64 inertia of the returned labels
and centroids (src/kmeans_lloyd_driver.hpp:356-418)."""
import numpy as np
import pytest
from numpy.testing import assert_array_equal

from oracle import c_oracle as co
from tests.conftest import check_label_mismatches, fp64_inertia, fp64_update, to_np

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import bench  # noqa: E402  (the workload generators of the benchmark)


def q():
    return torch.cuda.current_stream()


def drive(kd, X_d, w_d, init_d, iters, precision="fp32", tol=0.0):
    n = X_d.shape[1]
    lab = torch.empty(n, dtype=torch.int32, device=X_d.device)
    res = torch.empty_like(init_d)
    it, inertia, stats = kd.kmeans_lloyd_driver_ex(X_d, w_d, init_d.clone(), lab, res,
                                                   tol, False, iters, precision, q())
    torch.cuda.synchronize()
    return lab, res, int(it), float(inertia), stats


def rel_err(C, ref):
    ok = np.isfinite(ref).all(axis=0)
    return float(np.abs(C[:, ok] - ref[:, ok]).max() / np.abs(ref[:, ok]).max())


def test_c1_plain_fp32_full_size(kd, torch_dev):
    n, d, k = 10_000_000, 32, 256
    X_d, init_d = bench.make_data(n, d, k, bench.SEED, torch_dev)
    w_d = torch.ones(n, dtype=torch.float32, device=torch_dev)
    lab2, C2, it2, _, _ = drive(kd, X_d, w_d, init_d, 2)
    lab3, C3, it3, inr3, st = drive(kd, X_d, w_d, init_d, 3)
    assert (it2, it3) == (2, 3) and st["path"] == "fp32-mfma"
    Xt = to_np(X_d)
    w = np.ones(n)
    C2n, C3n, L2, L3 = to_np(C2), to_np(C3), to_np(lab2), to_np(lab3)
    # labels of the final centroids against the restatement
    m2 = check_label_mismatches(Xt, C2n, L2, co.assign(Xt, C2n), limit=n // 10_000)
    m3 = check_label_mismatches(Xt, C3n, L3, co.assign(Xt, C3n), limit=n // 10_000)
    print(f"C1 label mismatches vs restatement (near-ties): {m2}, {m3}")
    # iteration 3 is one update from the labels of C2
    ref, cnt = fp64_update(Xt, w, L2, k)
    assert rel_err(C3n, ref) <= 1e-5
    # the reference's fused step on C2 gives the same labels and exact counts
    hn = torch.empty(k, device=torch_dev)
    kd.half_l2_norm_squared(C2, hn, 256, q())
    ids = torch.empty(n, dtype=torch.int32, device=torch_dev)
    cp = torch.zeros((4, d, k), device=torch_dev)
    scp = torch.zeros((4, k), device=torch_dev)
    kd.fused_lloyd_single_step(X_d, w_d, C2, hn, ids, cp, scp, 8, 128, q())[0].wait()
    assert_array_equal(to_np(ids), L2)
    assert_array_equal(to_np(scp).astype(np.float64).sum(axis=0), cnt)
    sums = to_np(cp).astype(np.float64).sum(axis=0)
    assert np.abs(sums - ref * cnt).max() <= 1e-5 * np.abs(ref * cnt).max()
    # inertia of the returned labels and centroids
    assert abs(inr3 - fp64_inertia(Xt, w, C3n, L3)) <= 1e-5 * abs(inr3)


@pytest.mark.parametrize("k", [8, 32])
def test_few_centroids_full_size(kd, torch_dev, k):
    """n = 10,000,000, d = 32 with k = 8 and 32 (the few-centroid kernel of
    the HBM-bound regime), pinned like the k = 256 fit above; a BF16 request
    runs the same kernel and returns the same bits."""
    n, d = 10_000_000, 32
    X_d, init_d = bench.make_data(n, d, k, bench.SEED, torch_dev)
    w_d = torch.ones(n, dtype=torch.float32, device=torch_dev)
    lab2, C2, it2, _, _ = drive(kd, X_d, w_d, init_d, 2)
    lab3, C3, it3, inr3, st = drive(kd, X_d, w_d, init_d, 3)
    assert (it2, it3) == (2, 3) and st["path"] == "fp32-mfma"
    labb, Cb, _, inrb, stb = drive(kd, X_d, w_d, init_d, 3, "bf16")
    assert stb["path"] == "fp32-mfma"
    assert torch.equal(labb, lab3) and torch.equal(Cb.view(torch.int32), C3.view(torch.int32))
    assert inrb == inr3
    Xt = to_np(X_d)
    w = np.ones(n)
    C2n, C3n, L2, L3 = to_np(C2), to_np(C3), to_np(lab2), to_np(lab3)
    m2 = check_label_mismatches(Xt, C2n, L2, co.assign(Xt, C2n), limit=n // 10_000)
    m3 = check_label_mismatches(Xt, C3n, L3, co.assign(Xt, C3n), limit=n // 10_000)
    print(f"k={k} label mismatches vs restatement (near-ties): {m2}, {m3}")
    ref, _ = fp64_update(Xt, w, L2, k)
    assert rel_err(C3n, ref) <= 1e-5
    assert abs(inr3 - fp64_inertia(Xt, w, C3n, L3)) <= 1e-5 * abs(inr3)


@pytest.mark.parametrize("precision", ["bf16", "bf16x3", "fp8"])
def test_c2_mixed_precision_full_size(kd, torch_dev, precision):
    n, d, k = 2_000_000, 128, 1024
    X_d, init_d = bench.make_c2_data(n, d, k, 2001, torch_dev)
    w_d = torch.ones(n, dtype=torch.float32, device=torch_dev)
    labf, Cf, _, inrf, stf = drive(kd, X_d, w_d, init_d, 3, "fp32")
    labm, Cm, _, inrm, stm = drive(kd, X_d, w_d, init_d, 3, precision)
    assert stf["path"].startswith("tiled") and stm["path"].startswith("tiled")
    # the mixed form's labels and centroids are the FP32 form's, bit for bit
    assert torch.equal(labf, labm)
    assert torch.equal(Cf.view(torch.int32), Cm.view(torch.int32))
    assert inrm == inrf
    lab2, C2, _, _, _ = drive(kd, X_d, w_d, init_d, 2, precision)
    Xt = to_np(X_d)
    w = np.ones(n)
    C2n, C3n, L2, L3 = to_np(C2), to_np(Cm), to_np(lab2), to_np(labm)
    m = check_label_mismatches(Xt, C3n, L3, co.assign(Xt, C3n), limit=n // 10_000)
    print(f"C2 {precision}: label mismatches vs restatement (near-ties): {m}; "
          f"re-scored {stm['n_rescored'] / 3:.0f} samples per iteration")
    ref, _ = fp64_update(Xt, w, L2, k)
    err = rel_err(C3n, ref)
    inr_err = abs(inrm - fp64_inertia(Xt, w, C3n, L3)) / abs(inrm)
    print(f"C2 {precision}: centroid error vs FP64 update {err:.2e}, "
          f"inertia error vs FP64 {inr_err:.2e}")
    assert err <= 1e-5
    assert inr_err <= 1e-5


def _batch_args(data, dev):
    out = []
    for X_t, w, init in data:
        out.append((X_t, w, init.clone(),
                    torch.empty(X_t.shape[1], dtype=torch.int32, device=dev),
                    torch.empty_like(init), 1e-4, 100))
    return out


def test_c3_512_batched_fits_full_size(kd, torch_dev):
    data = bench.make_c3_batch(512, 3001, torch_dev)
    fits = _batch_args(data, torch_dev)
    res, _ = kd.kmeans_lloyd_batched(fits, "fp32", q(), 4)
    torch.cuda.synchronize()
    # every fit bitwise equal to the single-fit driver
    for i, ((X_t, w, init), (_, _, _, lab, out, tol, mi), (it_b, inr_b)) in \
            enumerate(zip(data, fits, res)):
        lab_s, out_s, it_s, inr_s, _ = drive(kd, X_t, w, init, mi, tol=tol)
        assert it_s == it_b, i
        assert torch.equal(lab, lab_s), i
        assert torch.equal(out.view(torch.int32), out_s.view(torch.int32)), i
        assert inr_s == inr_b, i
    # ten fits pinned step by step to the oracle: final labels against the
    # restatement, the last update against FP64, the stop rule, the inertia
    for i in range(0, 512, 52):
        X_t, w_d, init = data[i]
        _, _, _, lab, out, tol, mi = fits[i]
        it = int(res[i][0])
        Xt, w = to_np(X_t), to_np(w_d).astype(np.float64)
        k = init.shape[1]
        C, L = to_np(out), to_np(lab)
        check_label_mismatches(Xt, C, L, co.assign(Xt, C), limit=max(1, Xt.shape[1] // 10_000))
        if it >= 1:
            lab_p, C_p, _, _, _ = drive(kd, X_t, w_d, init, it - 1, tol=tol)
            ref, cnt = fp64_update(Xt, w, to_np(lab_p), k)
            if (cnt > 0).all():          # no relocation in the last step
                assert rel_err(C, ref) <= 1e-5, i
            shift = float(((C.astype(np.float64) - to_np(C_p)) ** 2).sum())
            assert shift <= tol * (1 + 1e-4) or it == mi, (i, shift)
        assert abs(float(res[i][1]) - fp64_inertia(Xt, w, C, L)) <= 1e-5 * abs(float(res[i][1]))


def test_c4_served_stream_with_large_requests(kd, torch_dev):
    from kmeans_dpcpp_amd import serving
    reqs = bench.make_c4_requests(24, 200.0, 4001, torch_dev)
    # two requests at the large size of the workload, first and in the middle
    for j, seed in ((0, 4101), (12, 4102)):
        X_t, init = bench.make_data(2_000_000, 96, 512, seed, torch_dev)
        reqs[j] = (reqs[j][0], X_t, torch.ones(2_000_000, device=torch_dev), init)
    torch.cuda.synchronize()
    budget = int(0.6 * torch.cuda.get_device_properties(0).total_memory)
    with serving.Server(budget, iterations_per_tick=1, large_precision="bf16") as srv:
        out = [srv.submit(X_t, w, init, tol=1e-4, max_iter=100) for _, X_t, w, init in reqs]
        for r in out:
            r.result(timeout=600)
    assert sum(not r.batched for r in out) >= 2
    for (_, X_t, w, init), r in zip(reqs, out):
        lab, C, it, inr, _ = drive(kd, X_t, w, init, 100, "fp32", tol=1e-4)
        assert it == r.n_iter
        assert torch.equal(lab, r.labels)
        assert torch.equal(C.view(torch.int32), r.centroids.view(torch.int32))
        assert inr == r.inertia
    # the large requests pinned to the oracle
    for (_, X_t, w_d, init), r in zip(reqs, out):
        if r.batched:
            continue
        Xt = to_np(X_t)
        w = np.ones(Xt.shape[1])
        C, L = to_np(r.centroids), to_np(r.labels)
        check_label_mismatches(Xt, C, L, co.assign(Xt, C), limit=Xt.shape[1] // 10_000)
        lab_p, _, _, _, _ = drive(kd, X_t, w_d, init, r.n_iter - 1, "bf16", tol=1e-4)
        ref, cnt = fp64_update(Xt, w, to_np(lab_p), init.shape[1])
        assert rel_err(C, ref) <= 1e-5
        assert abs(r.inertia - fp64_inertia(Xt, w, C, L)) <= 1e-5 * abs(r.inertia)
