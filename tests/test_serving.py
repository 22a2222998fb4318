"""Host-only tests of the serving loop's logic (kmeans_dpcpp_amd.serving):
admission by footprint against the budget, first-come-first-served order,
continuous batching (requests join a running set), the large-fit path,
latency bookkeeping and error reporting.  The GPU engine and the single-fit
runner are replaced by stand-ins; the GPU tests in test_gpu_serving.py run
the real ones."""
import threading
import time

import pytest

torch = pytest.importorskip("torch")

from kmeans_dpcpp_amd import serving  # noqa: E402


class FakeEngine:
    """Each fit 'runs' for max_iter steps; results mark the fit's order."""

    def __init__(self):
        self.fits = {}
        self.next = 0
        self.finished = []
        self.sets = []          # sizes of the running set at each step
        self.lock = threading.Lock()

    def add(self, X, w, init, labels, res, tol, max_iter):
        fid = self.next
        self.next += 1
        self.fits[fid] = [max_iter, 0, labels, res, init]
        return fid

    def step(self, iterations, use_graph):
        self.sets.append(len(self.fits))
        done = []
        for fid, f in self.fits.items():
            f[1] += iterations
            if f[1] >= f[0]:
                f[2].fill_(fid)
                f[3].copy_(f[4])
                done.append(fid)
        for fid in done:
            it = self.fits.pop(fid)[0]
            self.finished.append((fid, it, float(fid)))
        time.sleep(0.001)
        return len(self.fits)

    def take_finished(self):
        out, self.finished = self.finished, []
        return out


def make(n, d, k, seed=0):
    g = torch.Generator().manual_seed(seed)
    return (torch.randn(d, n, generator=g), torch.ones(n), torch.randn(d, k, generator=g))


def test_admission_respects_budget_and_order():
    budget = 10_000
    fp = lambda n, d, k: n * d  # noqa: E731
    eng = FakeEngine()
    with serving.Server(budget, engine_factory=lambda: eng, footprint=fp,
                        serves=lambda n, d, k: True) as srv:
        reqs = [srv.submit(*make(1000, 4, 3, i), max_iter=5 + i % 3) for i in range(12)]
        for r in reqs:
            r.result(timeout=10)
    assert srv.peak_resident <= budget
    # at most two 4000-byte requests fit at once
    assert max(eng.sets) <= 2
    admits = [r.t_admit for r in reqs]
    assert admits == sorted(admits)          # first come, first served
    assert all(r.latency is not None and r.latency >= 0 for r in reqs)


def test_requests_join_a_running_set():
    eng = FakeEngine()
    with serving.Server(1 << 30, engine_factory=lambda: eng,
                        footprint=lambda n, d, k: 1, serves=lambda n, d, k: True) as srv:
        first = srv.submit(*make(100, 2, 2), max_iter=400)
        while not eng.sets:
            time.sleep(0.001)
        late = srv.submit(*make(100, 2, 2, 1), max_iter=5)
        late.result(timeout=10)
        assert not first.done()                # the late one finished first
        first.result(timeout=10)
    assert 2 in eng.sets                       # both ran in one set


def test_large_requests_take_the_single_fit_path():
    ran = []

    def runner(req):
        ran.append(req.shape)
        req.labels = torch.zeros(req.shape[0], dtype=torch.int32)
        req.centroids = req.init_centroids_t.clone()
        return 3, 1.5

    eng = FakeEngine()
    with serving.Server(1 << 30, engine_factory=lambda: eng, large_runner=runner,
                        footprint=lambda n, d, k: 1,
                        serves=serving.engine_serves) as srv:
        big = srv.submit(*make(300, 100, 4), max_iter=10)
        small = srv.submit(*make(300, 8, 4), max_iter=2)
        assert big.result(timeout=10)[2:] == (3, 1.5)
        small.result(timeout=10)
    assert ran == [(300, 100, 4)]
    assert not big.batched and small.batched


def test_oversized_request_is_rejected_and_errors_are_reported():
    eng = FakeEngine()

    def runner(req):
        raise RuntimeError("device failure")

    with serving.Server(100, engine_factory=lambda: eng, large_runner=runner,
                        footprint=lambda n, d, k: n,
                        serves=lambda n, d, k: d < 50) as srv:
        with pytest.raises(ValueError):
            srv.submit(*make(1000, 2, 2))
        bad = srv.submit(*make(10, 60, 2))
        with pytest.raises(RuntimeError):
            bad.result(timeout=10)


def test_latency_summary():
    class R:
        def __init__(self, lat):
            self.latency = lat
    s = serving.latency_summary([R(x / 100) for x in range(1, 101)])
    assert s["count"] == 100
    assert abs(s["p50"] - 0.5) <= 0.011 and abs(s["p99"] - 0.99) <= 0.011


class FailingEngine(FakeEngine):
    """Steps fine until `fail` is set, then raises (as a HIP error would)."""

    def __init__(self, fail):
        super().__init__()
        self.fail = fail

    def step(self, iterations, use_graph):
        if self.fail.is_set():
            raise RuntimeError("hipErrorLaunchFailure (stand-in)")
        return super().step(iterations, use_graph)


def test_engine_error_finishes_every_request_and_closes_submit():
    fp = lambda n, d, k: n * d  # noqa: E731
    fail = threading.Event()
    srv = serving.Server(10**9, engine_factory=lambda: FailingEngine(fail),
                         footprint=fp, serves=lambda n, d, k: True)
    reqs = [srv.submit(*make(64, 4, 3, seed=i), max_iter=10**6) for i in range(6)]
    fail.set()
    for r in reqs:
        with pytest.raises(RuntimeError):
            r.result(timeout=10)
        assert r.done()
    with pytest.raises(RuntimeError):
        srv.submit(*make(64, 4, 3), max_iter=1)
    srv.close()


def test_engine_creation_error_fails_requests():
    def boom():
        raise RuntimeError("no device (stand-in)")
    srv = serving.Server(10**9, engine_factory=boom, footprint=lambda n, d, k: 1,
                         serves=lambda n, d, k: True)
    srv._loop.join(timeout=10)
    with pytest.raises(RuntimeError):
        srv.submit(*make(64, 4, 3), max_iter=1)
    srv.close()


def test_engine_serves_matches_native_limits():
    assert serving.engine_serves(1000, 64, 256)
    assert not serving.engine_serves(1000, 65, 8)
    assert not serving.engine_serves(1000, 16, 257)
    assert not serving.engine_serves(100, 16, 101)
    assert not serving.engine_serves(1 << 30, 16, 8)


class WideRangeEngine(FakeEngine):
    """Refuses fits of d == 5 as the native engine refuses data whose range
    needs floating-point sums."""

    def add(self, X, w, init, labels, res, tol, max_iter):
        if X.shape[0] == 5:
            raise ValueError("batched fit: the range of X or sample_weight needs "
                             "floating-point sums; use the single-fit driver")
        return super().add(X, w, init, labels, res, tol, max_iter)


def test_fits_the_engine_refuses_for_range_go_to_the_single_fit_driver():
    ran = []

    def runner(req):
        ran.append(req.shape)
        req.labels = torch.zeros(req.shape[0], dtype=torch.int32)
        req.centroids = req.init_centroids_t.clone()
        return 4, 2.5

    eng = WideRangeEngine()
    with serving.Server(1 << 30, engine_factory=lambda: eng, large_runner=runner,
                        footprint=lambda n, d, k: 1,
                        serves=lambda n, d, k: True) as srv:
        wide = srv.submit(*make(200, 5, 3), max_iter=3)
        plain = srv.submit(*make(200, 6, 3), max_iter=3)
        assert wide.result(timeout=10)[2:] == (4, 2.5)
        plain.result(timeout=10)
    assert ran == [(200, 5, 3)]
    assert not wide.batched and plain.batched
