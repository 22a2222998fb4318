"""GPU tests of continuous batching: fits that join a running engine set at
different times, and requests served by the request loop under a memory
budget (so some wait for admission) with a large fit running beside the
batch on its own stream -- every result equals what the single-fit driver
gives the same fit alone, bit for bit."""
import time

import numpy as np
import pytest
from numpy.testing import assert_array_equal

from oracle import lloyd_oracle as lo
from tests.conftest import to_dev, to_np

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def problem(n, d, k, seed):
    rng = np.random.default_rng(seed)
    Xt, init = lo.gaussian_blobs(n, d, k, seed=seed, std=float(rng.uniform(0.5, 2.0)))
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    return to_dev(Xt), to_dev(w), to_dev(init)


def alone(kd, X, w, init, tol, max_iter, precision="fp32"):
    lab = torch.empty(X.shape[1], dtype=torch.int32, device=X.device)
    res = torch.empty_like(init)
    it, inertia, _ = kd.kmeans_lloyd_driver_ex(X, w, init.clone(), lab, res, tol, False,
                                               max_iter, precision,
                                               torch.cuda.current_stream())
    torch.cuda.synchronize()
    return to_np(lab), to_np(res), int(it), float(inertia)


def same(got, ref):
    assert got[2] == ref[2]
    assert_array_equal(got[0], ref[0])
    assert_array_equal(got[1].view(np.uint32), ref[1].view(np.uint32))
    assert got[3] == ref[3]


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_engine_fits_joining_mid_run_equal_alone(kd, torch_dev, precision):
    probs = [problem(3000 + 500 * i, [16, 32, 64][i % 3], [8, 32, 64][i % 3], 40 + i)
             for i in range(6)]
    eng = kd.Engine(torch.cuda.current_stream(), precision)
    outs = {}

    def add(i):
        X, w, init = probs[i]
        lab = torch.empty(X.shape[1], dtype=torch.int32, device=X.device)
        res = torch.empty_like(init)
        work = init.clone()
        fid = eng.add(X, w, work, lab, res, 1e-4, 40)
        outs[fid] = (i, lab, res, work)

    for i in range(3):
        add(i)
    eng.step(3)
    for i in range(3, 6):          # join a set that is three iterations in
        add(i)
    results = {}
    while len(results) < 6:
        eng.step(2, use_graph=True)
        for fid, it, inertia in eng.take_finished():
            i, lab, res, _ = outs[fid]
            results[i] = (to_np(lab), to_np(res), int(it), float(inertia))
    for i, (X, w, init) in enumerate(probs):
        same(results[i], alone(kd, X, w, init, 1e-4, 40))


@pytest.mark.parametrize("small_precision", ["fp32", "bf16"])
def test_server_results_equal_alone_under_a_budget(kd, torch_dev, small_precision):
    from kmeans_dpcpp_amd import serving
    probs = [problem(2000 + 700 * i, [16, 32, 64, 8][i % 4], [8, 32, 64, 16][i % 4],
                     70 + i) for i in range(10)]
    big = problem(20_000, 100, 40, 99)
    foot = max(serving.request_footprint(X.shape[1], X.shape[0], init.shape[1])
               for X, _, init in probs)
    budget = 3 * foot + serving.request_footprint(20_000, 100, 40)
    with serving.Server(budget, iterations_per_tick=2,
                        small_precision=small_precision) as srv:
        reqs = []
        for i, (X, w, init) in enumerate(probs):
            reqs.append(srv.submit(X, w, init, tol=1e-4, max_iter=50))
            if i == 4:
                reqs.append(srv.submit(*big, tol=1e-4, max_iter=30))
            time.sleep(0.002)
        for r in reqs:
            r.result(timeout=120)
    assert srv.peak_resident <= budget
    expected = probs[:5] + [big] + probs[5:]
    for r, (X, w, init) in zip(reqs, expected):
        mi = 30 if r.shape[1] == 100 else 50
        same((to_np(r.labels), to_np(r.centroids), r.n_iter, r.inertia),
             alone(kd, X, w, init, 1e-4, mi))
    assert sum(1 for r in reqs if not r.batched) == 1
