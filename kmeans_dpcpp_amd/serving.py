"""Request loop for Lloyd fits on one GPU, with continuous batching.

The reference fits one problem per ``kmeans_lloyd_driver`` call
(src/kmeans_lloyd_driver.hpp:21-431); this module serves a stream of fit
requests on top of the same arithmetic:

* Admission by device footprint.  A request is admitted when the bytes the
  server holds for admitted, unfinished requests plus the request's own
  footprint fit the budget (first come, first served; a request that does not
  fit waits, and so does everything behind it).
* Continuous batching.  Admitted requests that the batch engine serves
  (d <= 64, k <= 256) join its running set at the next tick, whatever else
  is in flight; every tick advances all of them by ``iterations_per_tick``
  Lloyd iterations and retires the ones that stopped.  Their samples and
  centroid state stay on the device from admission to completion.
* Larger requests run through the single-fit driver on worker threads, each
  with a stream of its own, so their work overlaps the batch ticks (and each
  other) on the GPU.

Each request's result (labels, centroids, iteration count, inertia) is the
one ``kmeans_lloyd_driver`` gives it alone (the engine's results are
independent of the set it runs in), and the request records its submit,
admit and completion times for latency statistics.
"""
from __future__ import annotations

import collections
import threading
import time

from . import _kmeans_dpcpp_amd as _native


class FitRequest:
    """One fit: inputs, outputs (filled when done) and timestamps."""

    def __init__(self, X_t, sample_weight, init_centroids_t, tol, max_iter,
                 precision):
        self.X_t = X_t
        self.sample_weight = sample_weight
        self.init_centroids_t = init_centroids_t
        self.tol = float(tol)
        self.max_iter = int(max_iter)
        self.precision = precision
        self.labels = None
        self.centroids = None
        self.n_iter = None
        self.inertia = None
        self.error = None
        self.batched = None
        self.footprint = 0
        self.t_submit = time.perf_counter()
        self.t_admit = None
        self.t_done = None
        self._done = threading.Event()

    @property
    def shape(self):
        d, n = self.X_t.shape
        return n, d, self.init_centroids_t.shape[1]

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_submit

    def done(self) -> bool:
        return self._done.is_set()

    def result(self, timeout=None):
        if not self._done.wait(timeout):
            raise TimeoutError("fit request not finished")
        if self.error is not None:
            raise self.error
        return self.labels, self.centroids, self.n_iter, self.inertia


def engine_serves(n, d, k) -> bool:
    """The shapes the batch engine's kernels take (the native
    batched-fit check: d <= 64, k <= 256, k <= n < 2^30)."""
    return 1 <= d <= 64 and 1 <= k <= 256 and k <= n < (1 << 30)


def request_footprint(n, d, k) -> int:
    """Device bytes the server holds for a request while it runs: the engine
    workspace (or the single-fit driver's, with a sample-major copy of X on
    the tiled path), labels, result and working centroids."""
    own = 4 * n + 8 * d * k
    if engine_serves(n, d, k):
        return _native.Engine.footprint(n, d, k) + own
    return own + 4 * n * d + 16 * n + 8 * (d + 1) * k * 2 + (1 << 20)


class Server:
    """Serve fit requests with continuous batching on one GPU.

    ``submit`` may be called from any thread; a loop thread admits requests
    (first come first served, by footprint against ``hbm_budget_bytes``),
    ticks the batch engine and hands large fits to ``large_workers`` worker
    threads.
    ``small_precision`` is the batch engine's assignment ("fp32", or "bf16":
    the cross term in BF16 with a certified FP32 re-score, the same
    results); ``large_precision`` that of the single-fit driver runs.
    ``engine_factory`` / ``large_runner`` replace the GPU engine and the
    single-fit runner (for host-only tests of the loop's logic)."""

    def __init__(self, hbm_budget_bytes, iterations_per_tick=1, use_graph=False,
                 large_precision="fp32", engine_factory=None, large_runner=None,
                 footprint=request_footprint, serves=engine_serves, large_workers=2,
                 small_precision="fp32"):
        self.budget = int(hbm_budget_bytes)
        self.small_precision = small_precision
        self.iterations_per_tick = int(iterations_per_tick)
        self.use_graph = bool(use_graph)
        self.large_precision = large_precision
        self._engine_factory = engine_factory
        self._large_runner = large_runner or self._run_large
        self._footprint = footprint
        self._serves = serves
        self._lock = threading.Condition()
        self._pending = collections.deque()
        self._large = collections.deque()
        self._resident = 0
        self.peak_resident = 0
        self.ticks = 0
        self._running = True
        self._dead = None           # the error that ended the loop thread
        self._inflight = {}
        self._local = threading.local()
        self._loop = threading.Thread(target=self._run_loop, name="kmd-serve",
                                      daemon=True)
        self._workers = [threading.Thread(target=self._run_worker,
                                          name=f"kmd-serve-large-{i}", daemon=True)
                         for i in range(max(1, int(large_workers)))]
        self._loop.start()
        for t in self._workers:
            t.start()

    # ---- client side ------------------------------------------------------
    def submit(self, X_t, sample_weight, init_centroids_t, tol=1e-4,
               max_iter=100) -> FitRequest:
        req = FitRequest(X_t, sample_weight, init_centroids_t, tol, max_iter,
                         self.large_precision)
        n, d, k = req.shape
        req.batched = self._serves(n, d, k)
        req.footprint = int(self._footprint(n, d, k))
        if req.footprint > self.budget:
            raise ValueError("request needs more device memory than the server's "
                             "budget")
        with self._lock:
            if self._dead is not None:
                raise RuntimeError("the server's loop has stopped on an error") \
                    from self._dead
            if not self._running:
                raise RuntimeError("the server is closed")
            self._pending.append(req)
            self._lock.notify_all()
        return req

    def close(self):
        with self._lock:
            self._running = False
            self._lock.notify_all()
        self._loop.join()
        for t in self._workers:
            t.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- loop ----------------------------------------------------------------
    def _admit(self, engine):
        """Admit pending requests in order while the budget allows."""
        admitted = []
        with self._lock:
            while self._pending and \
                    self._resident + self._pending[0].footprint <= self.budget:
                req = self._pending.popleft()
                self._resident += req.footprint
                self.peak_resident = max(self.peak_resident, self._resident)
                admitted.append(req)
        for req in admitted:
            req.t_admit = time.perf_counter()
            if req.batched:
                try:
                    self._add_to_engine(engine, req)
                    continue
                except ValueError as e:
                    if "single-fit driver" not in str(e):
                        self._finish(req, error=e)
                        continue
                    # the data's range needs floating-point centroid sums,
                    # which only the single-fit driver has
                    req.batched = False
                except Exception as e:  # noqa: BLE001 - reported on the request
                    self._finish(req, error=e)
                    continue
            if not req.batched:
                with self._lock:
                    self._large.append(req)
                    self._lock.notify_all()

    def _add_to_engine(self, engine, req):
        import torch
        X = req.X_t
        n, d, k = req.shape
        req.labels = torch.empty(n, dtype=torch.int32, device=X.device)
        req.centroids = torch.empty((d, k), dtype=X.dtype, device=X.device)
        work = req.init_centroids_t.clone()  # the fit overwrites its init
        fid = engine.add(X, req.sample_weight, work, req.labels, req.centroids,
                         req.tol, req.max_iter)
        self._inflight[fid] = (req, work)

    def _finish(self, req, n_iter=None, inertia=None, error=None):
        req.n_iter = n_iter
        req.inertia = inertia
        req.error = error
        req.t_done = time.perf_counter()
        with self._lock:
            self._resident -= req.footprint
            self._lock.notify_all()
        req._done.set()

    def _run_loop(self):
        engine = None
        try:
            if self._engine_factory is not None:
                engine = self._engine_factory()
            else:
                engine = _native.Engine(None, self.small_precision)
            while True:
                with self._lock:
                    while self._running and not self._pending and not self._inflight:
                        self._lock.wait(0.05)
                    if not self._running and not self._pending and not self._inflight:
                        break
                self._admit(engine)
                if self._inflight:
                    engine.step(self.iterations_per_tick, self.use_graph)
                    self.ticks += 1
                    for fid, n_iter, inertia in engine.take_finished():
                        req, _ = self._inflight.pop(fid)
                        self._finish(req, n_iter, inertia)
                else:
                    with self._lock:
                        if self._pending and not self._inflight:
                            # waiting for memory held by large fits
                            self._lock.wait(0.005)
        except BaseException as e:  # noqa: BLE001 - every waiting request gets it
            self._fail_all(e)
        finally:
            del engine

    def _fail_all(self, error):
        """The loop thread cannot go on: finish every request it holds or
        would admit with `error` and refuse new submissions."""
        with self._lock:
            self._dead = error
            self._running = False
            stranded = [req for req, _ in self._inflight.values()]
            self._inflight.clear()
            stranded += list(self._pending)
            self._pending.clear()
            self._lock.notify_all()
        for req in stranded:
            if not req.done():
                self._finish(req, error=RuntimeError(
                    f"fit request abandoned: the server's loop stopped on "
                    f"{type(error).__name__}: {error}"))

    def _run_worker(self):
        while True:
            with self._lock:
                while self._running and not self._large:
                    self._lock.wait(0.05)
                if not self._large:
                    if not self._running:
                        return
                    continue
                req = self._large.popleft()
            try:
                n_iter, inertia = self._large_runner(req)
                self._finish(req, n_iter, inertia)
            except Exception as e:  # noqa: BLE001 - reported on the request
                self._finish(req, error=e)

    def _run_large(self, req):
        import torch
        from . import Stream
        if not hasattr(self._local, "stream"):  # one stream per worker thread
            self._local.stream = Stream()
        stream = self._local.stream
        X = req.X_t
        n, d, k = req.shape
        req.labels = torch.empty(n, dtype=torch.int32, device=X.device)
        req.centroids = torch.empty((d, k), dtype=X.dtype, device=X.device)
        work = req.init_centroids_t.clone()
        torch.cuda.current_stream().synchronize()  # clone done before the driver
        n_iter, inertia, _ = _native.kmeans_lloyd_driver_ex(
            X, req.sample_weight, work, req.labels, req.centroids, req.tol, False,
            req.max_iter, req.precision, stream.handle)
        stream.synchronize()
        return int(n_iter), float(inertia)


def latency_summary(requests):
    """p50 / p99 / mean latency (seconds) of finished requests."""
    lat = sorted(r.latency for r in requests if r.latency is not None)
    if not lat:
        return {"p50": None, "p99": None, "mean": None, "count": 0}

    def pct(p):
        i = min(len(lat) - 1, max(0, int(round(p / 100.0 * (len(lat) - 1)))))
        return lat[i]
    return {"p50": pct(50), "p99": pct(99), "mean": sum(lat) / len(lat),
            "count": len(lat)}
