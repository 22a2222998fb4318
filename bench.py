"""Benchmark of the Lloyd hot path on one MI355X.

Workload (the configuration the metric is quoted on): plain Lloyd, FP32,
n=10,000,000 samples, d=32, k=256, 256 Gaussian blobs with per-blob std
uniform in [0.5, 2.0] and centers uniform in [-20, 20]^32, seeded
random-sample init, tol=0.  A "step" is one Lloyd iteration (half-norms,
fused assignment+update, reduction, division, shifts) over the full sample
set resident in HBM.  The timed region is one kmeans_lloyd_driver_ex call
with max_iter = --steps (tol = 0, so exactly that many iterations run, plus
the driver's epilogue: final labels and inertia), bracketed by device
synchronisations.

Output: one JSON line with the metric, the roofline of the dominant kernel
(the fused assignment+update step, timed with HIP events inside the timed
region on the driver's stream) and a CPU baseline (the oracle's C
restatement on a bounded sample of the same workload, all host cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

METRIC = "Lloyd iters/s + % HBM/MFMA roofline; served fits/s with p50/p99 latency"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
FP64_PEAK_TFLOPS = 78.6

N, D, K = 10_000_000, 32, 256
SEED = 1001


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--precision", default="fp32", choices=["fp32", "bf16x3", "bf16", "fp8"])
    p.add_argument("--dtype", default="f32", choices=["f32", "f64"],
                   help="c1: element type of X, weights and centroids (f64: "
                        "the reference's double-precision entry points)")
    p.add_argument("--n", type=int, default=N)
    p.add_argument("--d", type=int, default=D)
    p.add_argument("--k", type=int, default=K)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--workload", default="c1", choices=["c1", "c2", "c3", "c4", "fit"],
                   help="c1: one plain fit n=10M d=32 k=256 (the metric's "
                        "configuration, default); c2: one fit n=2M d=128 "
                        "k=1024 on anisotropic blobs, 30 iterations, with the "
                        "BF16 cross term + exact re-score (--precision, default "
                        "bf16) beside plain FP32; c3: 512 independent fits in "
                        "one batched call (fits/s), with the same fits run one "
                        "by one through the single-fit driver beside it; c4: "
                        "a Poisson stream of fit requests served with "
                        "continuous batching (fits/s, p50/p99 latency) beside "
                        "the plain per-request FP32 driver; fit: the "
                        "sklearn-style estimator on sample-major data of the "
                        "c1 shape (column moments + centred transpose, GPU "
                        "k-means++ seeding, --steps Lloyd iterations), with "
                        "scikit-learn's k-means++ on a host sample beside it")
    p.add_argument("--rate", type=float, default=200.0, help="c4 arrivals per s")
    p.add_argument("--large-precision", default="bf16", choices=["fp32", "bf16x3", "bf16"],
                   help="c4: precision of the large requests' fits")
    p.add_argument("--small-precision", default="bf16", choices=["fp32", "bf16"],
                   help="c4: assignment precision of the batch engine")
    p.add_argument("--no-extras", action="store_true",
                   help="c1: leave out the batched-fit, served-stream and BF16 "
                        "measurements that the default line carries beside the "
                        "FP32 figure")
    p.add_argument("--extra-requests", type=int, default=200,
                   help="c1 extras: requests of the served stream")
    p.add_argument("--saturated-rate", type=float, default=2000.0,
                   help="c1 extras: arrival rate of the second served stream "
                        "(above what either form serves)")
    p.add_argument("--requests", type=int, default=400, help="c4 requests")
    p.add_argument("--fits", type=int, default=512)
    p.add_argument("--copy-probe-mib", type=int, default=0,
                   help="first run one device-to-device copy of this many MiB "
                        "(a kernel of known HBM traffic: calibrates rocprofv3 "
                        "FETCH_SIZE / WRITE_SIZE)")
    return p.parse_args()


def make_data(n, d, k, seed, device):
    """Blobs generated on the GPU in the reference's transposed layout."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    centers = torch.rand((k, d), generator=g, device=device) * 40.0 - 20.0
    std = torch.rand((k,), generator=g, device=device) * 1.5 + 0.5
    blob = torch.randint(0, k, (n,), generator=g, device=device)
    X_t = torch.empty((d, n), dtype=torch.float32, device=device)
    chunk = 1 << 22
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        b = blob[s:e]
        z = torch.randn((e - s, d), generator=g, device=device)
        X_t[:, s:e] = (centers[b] + z * std[b, None]).T
    init_idx = torch.randperm(n, generator=g, device=device)[:k]
    init_t = X_t[:, init_idx].contiguous()
    return X_t, init_t


def make_c2_data(n, d, k, seed, device):
    """Anisotropic Gaussian blobs: blob b draws x = c_b + Q_b (s_b * z) with
    z ~ N(0, I), Q_b a random orthogonal matrix (QR of a Gaussian matrix,
    signs fixed by diag(R) > 0), axis scales s_b log-uniform in [0.1, 3] and
    centers c_b uniform in [-20, 20]^d; equal-probability blob choice; seeded
    random-sample init.  Returns X_t (d, n) and the init (d, k)."""
    import math
    import torch
    gc = torch.Generator()
    gc.manual_seed(seed)
    a = torch.randn((k, d, d), generator=gc, dtype=torch.float64)
    q, r = torch.linalg.qr(a)
    q = q * torch.sign(torch.diagonal(r, dim1=-2, dim2=-1))[:, None, :]
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    qt = q.transpose(1, 2).contiguous().to(device=device, dtype=torch.float32)
    scales = torch.exp(torch.rand((k, d), generator=g, device=device)
                       * (math.log(3.0) - math.log(0.1)) + math.log(0.1))
    centers = torch.rand((k, d), generator=g, device=device) * 40.0 - 20.0
    blob = torch.randint(0, k, (n,), generator=g, device=device)
    order = torch.argsort(blob, stable=True)
    sb = blob[order]
    counts = torch.bincount(sb, minlength=k)
    start = torch.cumsum(counts, 0) - counts
    rank = torch.arange(n, device=device) - start[sb]
    m = int(counts.max())
    z = torch.randn((k, m, d), generator=g, device=device) * scales[:, None, :]
    y = torch.bmm(z, qt)                      # row j of blob b: (Q_b (s_b z))^T
    del z
    rows = torch.empty((n, d), dtype=torch.float32, device=device)
    rows[order] = y[sb, rank] + centers[sb]
    del y
    X_t = rows.T.contiguous()
    del rows
    init_idx = torch.randperm(n, generator=g, device=device)[:k]
    return X_t, X_t[:, init_idx].contiguous()


def fp64_update(X_t, w, labels, k):
    """One FP64 centroid update from given labels (empty clusters: NaN):
    samples sorted by label, then per-cluster FP64 segment sums (no atomics,
    so its cost does not grow with few, crowded clusters)."""
    import torch
    lab = labels.long()
    order = torch.argsort(lab, stable=True)
    lens = torch.bincount(lab, minlength=k)
    ws = w[order].double()
    cnt = torch.segment_reduce(ws, "sum", lengths=lens, axis=0)
    xs = X_t.T[order].double() * ws[:, None]
    sums = torch.segment_reduce(xs, "sum", lengths=lens, axis=0)
    del xs
    return (sums / cnt[:, None]).T, cnt


def fp64_inertia(X_t, w, labels, C_t):
    """FP64 sum of w * ||x - c_label||^2 for the given labels and centroids."""
    import torch
    C64 = C_t.double()
    lab = labels.long()
    tot = torch.zeros((), dtype=torch.float64, device=X_t.device)
    chunk = 1 << 20
    for s in range(0, X_t.shape[1], chunk):
        e = min(X_t.shape[1], s + chunk)
        diff = X_t[:, s:e].double() - C64[:, lab[s:e]]
        tot += (w[s:e].double() * (diff * diff).sum(0)).sum()
    return float(tot)


def fp64_errors(X_t, w, labels_prev, centroids, labels, inertia):
    """Error of one fit's result against FP64: the centroids against one FP64
    update from the labels of the iteration before the last (clusters empty
    in those labels left out: they were relocated), the inertia against the
    FP64 inertia of the final labels and centroids."""
    k = centroids.shape[1]
    c64, cnt = fp64_update(X_t, w, labels_prev, k)
    ok = cnt > 0
    c_err = 0.0
    if bool(ok.any()):
        c_err = float((centroids.double() - c64)[:, ok].abs().max()
                      / c64[:, ok].abs().max().clamp_min(1e-300))
    i64 = fp64_inertia(X_t, w, labels, centroids)
    i_err = abs(float(inertia) - i64) / max(abs(i64), 1e-300)
    return {"centroid_rel_err": c_err, "inertia_rel_err": i_err,
            "empty_before_last_update": int((~ok).sum())}


def sampled_fp64_errors(fits, stream):
    """fp64_errors of finished fits, given as (X_t, w, init, n_iter, centroids,
    labels, inertia): the labels of the iteration before the last come from
    the plain FP32 driver rerun for n_iter - 1 iterations from the same init
    (batched, served and mixed-precision fits equal the plain driver's bit for
    bit).  Returns the largest errors over the fits and how many were
    checked."""
    import torch
    import kmeans_dpcpp_amd as kd
    worst_c, worst_i, empty, checked = 0.0, 0.0, 0, 0
    for X_t, w, init, n_iter, centroids, labels, inertia in fits:
        if n_iter < 1:
            continue
        lab_prev = torch.empty_like(labels)
        c_a, c_b = init.clone(), torch.empty_like(init)
        kd.kmeans_lloyd_driver_ex(X_t, w, c_a, lab_prev, c_b, 0.0, False, int(n_iter) - 1,
                                  "fp32", stream)
        torch.cuda.synchronize()
        e = fp64_errors(X_t, w, lab_prev, centroids, labels, inertia)
        worst_c = max(worst_c, e["centroid_rel_err"])
        worst_i = max(worst_i, e["inertia_rel_err"])
        empty += e["empty_before_last_update"]
        checked += 1
    return {"fits_checked": checked, "centroid_rel_err_max": worst_c,
            "inertia_rel_err_max": worst_i, "empty_before_last_update": empty}


def bench_c2(args):
    """Compute-bound single fit: the mixed-precision driver against plain FP32
    on the same data, with an FP64 check of the last centroid update."""
    import torch
    import kmeans_dpcpp_amd as kd
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()
    n, d, k = 2_000_000, 128, 1024
    steps = args.steps if args.steps != 50 else 30
    prec = args.precision if args.precision != "fp32" else "bf16"
    X_t, init_t = make_c2_data(n, d, k, 2001, dev)
    w = torch.ones(n, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    def run(p, iters, warm):
        lab = torch.empty(n, dtype=torch.int32, device=dev)
        c_a, c_b = init_t.clone(), torch.empty_like(init_t)
        if warm:
            kd.kmeans_lloyd_driver_ex(X_t, w, c_a, lab, c_b, 0.0, False,
                                      args.warmup, p, stream)
            c_a = init_t.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it, inertia, stats = kd.kmeans_lloyd_driver_ex(X_t, w, c_a, lab, c_b, 0.0,
                                                      False, iters, p, stream)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        assert it == iters, (it, iters)
        return el, lab, c_b, inertia, stats

    el_m, lab_m, cen_m, inr_m, st_m = run(prec, steps, args.warmup > 0)
    el_p, lab_p, cen_p, inr_p, st_p = run("fp32", steps, args.warmup > 0)
    # FP64 check: centroids after `steps` iterations against one FP64 update
    # from the labels the mixed form gives for its centroids after steps - 1
    _, lab_prev, _, _, _ = run(prec, steps - 1, False)
    err = fp64_errors(X_t, w, lab_prev, cen_m, lab_m, inr_m)

    kern_ms = st_m["fused_kernel_ms"] / max(1, st_m["fused_launches"])
    flops = 2.0 * n * k * d + 2.0 * n * d
    tflops = flops / (kern_ms * 1e-3) / 1e12
    bytes_per_launch = n * d * 4 + n * 4 + n * 4
    roof = {"bound": "mfma", "achieved": round(tflops, 3), "peak": BF16_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(tflops / BF16_PEAK_TFLOPS, 4),
            "traffic": None,
            "kernel": "assignment (BF16 MFMA) + exact re-score + exact update",
            "kernel_ms": round(kern_ms, 4),
            "hbm_frac": round(bytes_per_launch / (kern_ms * 1e-3) / 1e9
                              / HBM_PEAK_GBS, 4),
            "algorithmic_bytes": bytes_per_launch, "algorithmic_flops": flops}
    kp = st_p["fused_kernel_ms"] / max(1, st_p["fused_launches"])
    line = {
        "metric": METRIC,
        "value": round(steps / el_m, 3),
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * el_m / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": prec,
        "data": "synthetic",
        "config": {"workload": f"mixed-precision Lloyd, n={n:,} d={d} k={k} "
                               f"anisotropic Gaussian blobs, tol=0, {steps} "
                               f"iterations",
                   "n_samples": n, "n_features": d, "n_clusters": k,
                   "precision": prec},
        "roofline": roof,
        "cpu_baseline": None,
        "path": st_m.get("path"),
        "rescored_per_iter": st_m.get("n_rescored", 0) / steps,
        "plain_fp32": {"iters_per_s": round(steps / el_p, 3),
                       "ms_per_step": round(1e3 * el_p / steps, 4),
                       "kernel_ms": round(kp, 4),
                       "frac_fp32_peak": round(flops / (kp * 1e-3) / 1e12
                                               / FP32_PEAK_TFLOPS, 4),
                       "path": st_p.get("path")},
        "speedup_vs_fp32": round(el_p / el_m, 2),
        "labels_identical_to_fp32": bool(torch.equal(lab_m, lab_p)),
        "centroids_identical_to_fp32": bool(torch.equal(cen_m.view(torch.int32),
                                                        cen_p.view(torch.int32))),
        "inertia": [inr_m, inr_p],
        "error_vs_fp64": err,
    }
    print(json.dumps(line), flush=True)


def make_c3_batch(n_fits, seed, device):
    """512 independent fits: n log-uniform in [1e3, 2e5], d in {16, 32, 64},
    k in {8, 32, 64, 128, 256} with k < n / 10; Gaussian blobs (centers
    uniform in [-20, 20]^d, per-blob std uniform in [0.5, 2]); seeded
    random-sample init; tol 1e-4, max_iter 100."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    fits = []
    for _ in range(n_fits):
        n = int(np.exp(rng.uniform(np.log(1e3), np.log(2e5))))
        d = int(rng.choice([16, 32, 64]))
        k = int(rng.choice([8, 32, 64, 128, 256]))
        while k >= n / 10:
            k = int(rng.choice([8, 32, 64, 128, 256]))
        centers = torch.rand((k, d), generator=g, device=device) * 40.0 - 20.0
        std = torch.rand((k,), generator=g, device=device) * 1.5 + 0.5
        blob = torch.randint(0, k, (n,), generator=g, device=device)
        z = torch.randn((n, d), generator=g, device=device)
        X_t = (centers[blob] + z * std[blob, None]).T.contiguous()
        init_idx = torch.randperm(n, generator=g, device=device)[:k]
        fits.append((X_t, torch.ones(n, device=device), X_t[:, init_idx].contiguous()))
    return fits


def run_c3(n_fits, precision):
    """n_fits independent fits in one batched call (assignment precision
    'fp32' or 'bf16'), then the same fits one by one through the plain FP32
    single-fit driver; every fit is compared bitwise."""
    import torch
    import kmeans_dpcpp_amd as kd
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()
    data = make_c3_batch(n_fits, 3001, dev)
    n_total = sum(x.shape[1] for x, _, _ in data)

    def batch_args():
        out = []
        for X_t, w, init in data:
            out.append((X_t, w, init.clone(),
                        torch.empty(X_t.shape[1], dtype=torch.int32, device=dev),
                        torch.empty_like(init), 1e-4, 100))
        return out

    kd.kmeans_lloyd_batched(batch_args()[:8], precision, stream)      # warm-up
    fits = batch_args()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res, stats = kd.kmeans_lloyd_batched(fits, precision, stream, 4)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    iters = [r[0] for r in res]
    # the same fits one by one through the single-fit driver
    seq = batch_args()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    seq_out = []
    for (X_t, w, init, lab, out, tol, mi) in seq:
        it_s, inr_s = kd.kmeans_lloyd_driver(X_t, w, init, lab, out, tol, False, mi,
                                             8, 128, 0.7, stream)
        seq_out.append((it_s, float(inr_s[0])))
    torch.cuda.synchronize()
    el_seq = time.perf_counter() - t1
    differ = []
    for i, ((X_t, _, _, lab, out, _, _), (it_s, inr_s), (it_b, inr_b), bf) in \
            enumerate(zip(seq, seq_out, res, fits)):
        # bitwise: an empty, unrelocated cluster is NaN in both
        if it_s != it_b or not torch.equal(lab, bf[3]) or \
                not torch.equal(out.view(torch.int32), bf[4].view(torch.int32)):
            differ.append({"fit": i, "n": int(X_t.shape[1]), "d": int(X_t.shape[0]),
                           "k": int(out.shape[1]), "iters": [int(it_s), int(it_b)],
                           "label_mismatch": int((lab != bf[3]).sum()),
                           "max_centroid_diff": float((out - bf[4]).abs().max())})
    same = not differ
    sample = range(0, len(fits), max(1, len(fits) // 8))
    err = sampled_fp64_errors(
        [(data[i][0], data[i][1], data[i][2], int(res[i][0]), fits[i][4], fits[i][3],
          float(res[i][1])) for i in sample], stream)
    return {
        "fits_per_s": round(len(fits) / el, 3),
        "ms": round(1e3 * el, 3),
        "precision": precision,
        "workload": f"{len(fits)} independent Lloyd fits in one batched call (n "
                    f"log-uniform 1e3..2e5, d in {{16,32,64}}, k in {{8..256}}, tol "
                    f"1e-4, max_iter 100)",
        "total_samples": n_total,
        "graph_launches": stats["graph_launches"],
        "iterations": {"sum": int(sum(iters)), "max": int(max(iters)),
                       "mean": round(sum(iters) / len(iters), 2)},
        "sequential_single_fit_fits_per_s": round(len(fits) / el_seq, 3),
        "speedup_vs_sequential": round(el_seq / el, 2),
        "identical_to_single_fit": bool(same),
        "fits_differing": differ[:5],
        "n_fits_differing": len(differ),
        "error_vs_fp64": err,
    }


def bench_c3(args):
    prec = args.precision if args.precision in ("fp32", "bf16") else "bf16"
    r = run_c3(args.fits, prec)
    line = {
        "metric": METRIC,
        "value": r["fits_per_s"],
        "unit": "fits/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": r["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == "fp32" else "bf16 cross term + f32 re-score",
        "data": "synthetic",
        "config": {"workload": r["workload"], "total_samples": r["total_samples"],
                   "precision": prec},
    }
    line.update({k: v for k, v in r.items()
                 if k not in ("fits_per_s", "ms", "workload", "total_samples")})
    print(json.dumps(line), flush=True)


def make_c4_requests(n_req, rate, seed, device):
    """Poisson arrivals; 95% of requests drawn like the batched fits above,
    5% large (n=2,000,000, d=96, k=512); per-request seeds."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_req))
    reqs = []
    for i in range(n_req):
        if rng.uniform() < 0.05:
            n, d, k = 2_000_000, 96, 512
        else:
            n = int(np.exp(rng.uniform(np.log(1e3), np.log(2e5))))
            d = int(rng.choice([16, 32, 64]))
            k = int(rng.choice([8, 32, 64, 128, 256]))
            while k >= n / 10:
                k = int(rng.choice([8, 32, 64, 128, 256]))
        g = torch.Generator(device=device)
        g.manual_seed(seed * 100_003 + i)
        centers = torch.rand((k, d), generator=g, device=device) * 40.0 - 20.0
        std = torch.rand((k,), generator=g, device=device) * 1.5 + 0.5
        blob = torch.randint(0, k, (n,), generator=g, device=device)
        X_t = torch.empty((d, n), device=device)
        for s0 in range(0, n, 1 << 20):
            e = min(n, s0 + (1 << 20))
            z = torch.randn((e - s0, d), generator=g, device=device)
            X_t[:, s0:e] = (centers[blob[s0:e]] + z * std[blob[s0:e], None]).T
        init = X_t[:, torch.randperm(n, generator=g, device=device)[:k]].contiguous()
        reqs.append((float(t[i]), X_t, torch.ones(n, device=device), init))
    return reqs


def run_c4(n_requests, rate, large_precision, small_precision):
    """A Poisson stream served with continuous batching, then the same
    arrivals served one at a time by the plain FP32 driver; every request's
    result is compared bitwise."""
    import torch
    import kmeans_dpcpp_amd as kd
    from kmeans_dpcpp_amd import serving
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()
    reqs = make_c4_requests(n_requests, rate, 4001, dev)
    torch.cuda.synchronize()
    tol, max_iter = 1e-4, 100

    # served: continuous batching; the large fits use the BF16 cross term +
    # exact re-score (labels and centroids identical to FP32) on their own
    # stream
    budget = int(0.6 * torch.cuda.get_device_properties(0).total_memory)
    out = []
    with serving.Server(budget, iterations_per_tick=1, large_precision=large_precision,
                        small_precision=small_precision) as srv:
        t0 = time.perf_counter()
        for (ta, X_t, w, init) in reqs:
            now = time.perf_counter() - t0
            if ta > now:
                time.sleep(ta - now)
            r = srv.submit(X_t, w, init, tol=tol, max_iter=max_iter)
            out.append(r)
        for r in out:
            r.result(timeout=600)
        el = max(r.t_done for r in out) - t0
    lat = serving.latency_summary(out)
    # plain per-request FP32 form: one fit at a time, first come first served
    plain_lat, same = [], True
    t0 = time.perf_counter()
    for (ta, X_t, w, init), r in zip(reqs, out):
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        lab = torch.empty(X_t.shape[1], dtype=torch.int32, device=dev)
        res = torch.empty_like(init)
        it, inertia = kd.kmeans_lloyd_driver(X_t, w, init.clone(), lab, res, tol, False,
                                             max_iter, 8, 128, 0.7, stream)
        torch.cuda.synchronize()
        plain_lat.append(time.perf_counter() - t0 - ta)
        same &= (it == r.n_iter) and bool(torch.equal(lab, r.labels)) and bool(
            torch.equal(res.view(torch.int32), r.centroids.view(torch.int32)))
    el_plain = time.perf_counter() - t0
    plain_lat.sort()
    small = [i for i, r in enumerate(out) if r.batched]
    pick = [i for i, r in enumerate(out) if not r.batched] + \
        small[::max(1, len(small) // 8)][:8]
    err = sampled_fp64_errors(
        [(reqs[i][1], reqs[i][2], reqs[i][3], out[i].n_iter, out[i].centroids,
          out[i].labels, out[i].inertia) for i in pick], stream)

    def pct(v, p):
        return v[min(len(v) - 1, int(round(p / 100 * (len(v) - 1))))]
    return {
        "fits_per_s": round(len(out) / el, 3),
        "ms": round(1e3 * el, 3),
        "workload": f"served stream: {len(out)} Poisson arrivals at {rate:g} req/s, "
                    f"95% batched-fit sizes, 5% n=2M d=96 k=512; tol 1e-4, max_iter "
                    f"100; continuous batching, footprint admission; small fits "
                    f"{small_precision}, large fits {large_precision}",
        "rate_req_per_s": rate,
        "requests": len(out),
        "large_requests": sum(1 for r in out if not r.batched),
        "latency_s": {"p50": round(lat["p50"], 4), "p99": round(lat["p99"], 4),
                      "mean": round(lat["mean"], 4)},
        "plain_per_request_fp32": {
            "fits_per_s": round(len(out) / el_plain, 3),
            "latency_s": {"p50": round(pct(plain_lat, 50), 4),
                          "p99": round(pct(plain_lat, 99), 4)}},
        "speedup_fits_per_s": round(el_plain / el, 2),
        "engine_ticks": srv.ticks,
        "peak_resident_gb": round(srv.peak_resident / 1e9, 2),
        "identical_to_plain": bool(same),
        "error_vs_fp64": err,
    }


def bench_c4(args):
    r = run_c4(args.requests, args.rate, args.large_precision, args.small_precision)
    line = {
        "metric": METRIC,
        "value": r["fits_per_s"],
        "unit": "fits/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": r["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 cross term + f32 re-score",
        "data": "synthetic",
        "config": {"workload": r["workload"]},
    }
    line.update({k: v for k, v in r.items() if k not in ("fits_per_s", "ms", "workload")})
    print(json.dumps(line), flush=True)


def copy_probe(mib, dev):
    """One device-to-device copy of `mib` MiB: a kernel of known HBM traffic
    that calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE."""
    if mib <= 0:
        return
    import torch
    src = torch.full((mib << 18,), 1.0, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def bench_fit(args):
    """The estimator front-end end to end: KMeans(k, init='k-means++',
    n_init=1, max_iter=steps, tol=0).fit(X) on sample-major float32 X of the
    c1 shape, with its parts timed separately (HIP events on the stream the
    kernels run on) and the k-means++ seeding priced against the HBM roof."""
    import numpy as np
    import torch
    import kmeans_dpcpp_amd as kd
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()
    n, d, k = args.n, args.d, args.k
    steps = args.steps
    X_t0, _ = make_data(n, d, k, SEED, dev)
    X = X_t0.t().contiguous()                      # (n, d) sample-major
    del X_t0
    copy_probe(args.copy_probe_mib, dev)
    w = torch.ones(n, dtype=torch.float32, device=dev)
    T = 2 + int(np.log(k))
    torch.cuda.synchronize()
    # warm-up: every kernel of the fit once on a slice
    if args.warmup > 0:
        kd.KMeans(k, random_state=0, max_iter=2, tol=0).fit(X[:max(4 * k, 200_000)])
        torch.cuda.synchronize()

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        out = fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1), out

    # parts, each on the full data
    mean = torch.empty(d, dtype=torch.float64, device=dev)
    var = torch.empty(d, dtype=torch.float64, device=dev)
    X_t = torch.empty((d, n), dtype=torch.float32, device=dev)

    def prep():
        kd.column_mean_var(X, mean, var, stream)
        kd.transpose(X, X_t, mean, stream)
    prep_ms, _ = timed(prep)
    u = np.random.RandomState(0).random_sample((k, T))
    C = torch.empty((d, k), dtype=torch.float32, device=dev)
    idx = torch.empty(k, dtype=torch.int64, device=dev)
    kpp_ms, _ = timed(lambda: kd.kmeans_plusplus_t(X_t, w, u, C, idx, stream))
    del X_t
    torch.cuda.synchronize()
    # the whole fit
    t0 = time.perf_counter()
    est = kd.KMeans(k, init="k-means++", n_init=1, max_iter=steps, tol=0,
                    random_state=0).fit(X)
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t0

    # k-means++ traffic: the first center reads w; the pass of center 0
    # reads X and w and writes D^2; every later pass reads X, w and D^2 and
    # writes D^2 (FP64)
    kpp_bytes = (n * 4 + n * (4 * d + 4 + 8)
                 + (k - 1) * n * (4 * d + 4 + 8 + 8))
    kpp_gbs = kpp_bytes / (kpp_ms * 1e-3) / 1e9
    here = os.path.dirname(os.path.abspath(__file__))
    # HBM bytes per scoring launch measured with rocprofv3 (profiled shape only)
    traffic = (traffic_from_profile(os.path.join(here, "profiles",
                                                 "kmeanspp_fit_n10M_d32_k256.md"))
               if (n, d, k) == (N, D, K) else None)
    roof = {"bound": "hbm", "achieved": round(kpp_gbs, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(kpp_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "k-means++ seeding (kpp_score_krn + one-block kpp_pick_krn per center)",
            "kernel_ms": round(kpp_ms, 3),
            "algorithmic_bytes": kpp_bytes}
    prep_bytes = 2 * n * d * 4 + 2 * n * d * 4      # two moment passes + transpose
    line = {
        "metric": METRIC,
        "value": round(1.0 / fit_s, 4),
        "unit": "fits/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * fit_s / max(1, est.n_iter_), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (seeding: f64)",
        "data": "synthetic",
        "config": {"workload": f"estimator fit: KMeans(n_clusters={k}, "
                               f"init='k-means++', n_init=1, max_iter={steps}, "
                               f"tol=0) on sample-major float32 blobs, "
                               f"n={n:,} d={d}",
                   "n_samples": n, "n_features": d, "n_clusters": k,
                   "n_local_trials": T},
        "roofline": roof,
        "fit_s": round(fit_s, 4),
        "n_iter": est.n_iter_,
        "parts_ms": {"moments_and_centred_transpose": round(prep_ms, 3),
                     "kmeans_plusplus": round(kpp_ms, 3),
                     "lloyd_and_rest": round(1e3 * fit_s - prep_ms - kpp_ms, 3)},
        "prep_hbm_frac": round(prep_bytes / (prep_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "cpu_baseline": None if args.no_cpu_baseline else
        cpu_kmeanspp_baseline(X, k, T, n, args.cpu_seconds),
    }
    if line["cpu_baseline"]:
        line["kmeanspp_gpu_over_cpu"] = round(
            (1e3 / kpp_ms) / line["cpu_baseline"]["value"], 1)
    print(json.dumps(line), flush=True)


def cpu_kmeanspp_baseline(X, k, T, n, budget_s):
    """scikit-learn's k-means++ (its own BLAS threads, all host cores) on a
    host sample of the same data, scaled linearly in n to the full set."""
    import numpy as np
    try:
        from sklearn.cluster import kmeans_plusplus
    except ImportError:
        return None
    cores = len(os.sched_getaffinity(0))
    n_s = min(n, 20_000)
    Xs = X[:n_s].cpu().numpy().copy()
    t0 = time.perf_counter()
    kmeans_plusplus(Xs, k, random_state=0, n_local_trials=T)
    el = time.perf_counter() - t0
    while el < budget_s / 4 and n_s < n:
        n_s = min(n, n_s * 4)
        Xs = X[:n_s].cpu().numpy().copy()
        t0 = time.perf_counter()
        kmeans_plusplus(Xs, k, random_state=0, n_local_trials=T)
        el = time.perf_counter() - t0
    full_s = el * (n / n_s)
    return {"value": round(1.0 / full_s, 5), "unit": "k-means++ seedings/s",
            "cores": cores, "kind": "port",
            "sample": (f"scikit-learn {__import__('sklearn').__version__} "
                       f"kmeans_plusplus (float32, k={k}, {T} local trials) on "
                       f"the first {n_s:,} samples in {el:.2f} s, scaled "
                       f"linearly to n={n:,}")}


def traffic_from_profile(path):
    """HBM bytes per fused-step launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration)."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


def cpu_baseline(X_t, init_t, n, d, k, budget_s):
    """The oracle's C restatement (OpenMP, all host cores) on a bounded sample
    of the same workload; Lloyd iterations/s scaled to the full n."""
    import numpy as np
    here = os.path.dirname(os.path.abspath(__file__))
    if here not in sys.path:
        sys.path.insert(0, here)
    from oracle import c_oracle
    cores = len(os.sched_getaffinity(0))
    threads = int(os.environ.get("OMP_NUM_THREADS", cores))
    threads = max(1, min(threads, cores))
    n_s = min(n, 2_000_000)
    Xs = X_t[:, :n_s].cpu().numpy().copy()
    Cs = init_t.cpu().numpy().copy()
    w = np.ones(n_s, dtype=np.float32)
    t0 = time.perf_counter()
    c_oracle.lloyd(Xs, w, Cs, 0.0, 1, nthreads=threads)
    t1 = time.perf_counter() - t0          # one iteration + epilogue (~2 passes)
    iters = max(1, min(500, int(budget_s / max(t1 / 2, 1e-3)) - 1))
    t0 = time.perf_counter()
    _, _, it, _ = c_oracle.lloyd(Xs, w, Cs, 0.0, iters, nthreads=threads)
    el = time.perf_counter() - t0
    per_iter_full = (el / (it + 1)) * (n / n_s)   # epilogue costs ~one pass
    return {
        "value": 1.0 / per_iter_full,
        "unit": "iters/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"C restatement of the reference (oracle/restate.c), FP32, "
                   f"{it} Lloyd iterations + final assignment on the first "
                   f"{n_s:,} of the {n:,} samples (d={d}, k={k}) in {el:.1f} s; "
                   f"scaled linearly to n={n:,}"),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the Lloyd "
              f"path is single-GPU in this build", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c2":
        bench_c2(args)
        return
    if args.workload == "c3":
        bench_c3(args)
        return
    if args.workload == "c4":
        bench_c4(args)
        return
    if args.workload == "fit":
        bench_fit(args)
        return
    import torch
    import kmeans_dpcpp_amd as kd

    dev = torch.device("cuda", 0)
    n, d, k = args.n, args.d, args.k
    X_t, init_t = make_data(n, d, k, SEED, dev)
    if args.dtype == "f64":
        if args.precision != "fp32":
            sys.exit("--dtype f64 runs the FP64 arithmetic only (--precision fp32)")
        X_t, init_t = X_t.double(), init_t.double()
    w = torch.ones(n, dtype=X_t.dtype, device=dev)
    labels = torch.empty(n, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    copy_probe(args.copy_probe_mib, dev)

    # warmup: W iterations (untimed)
    c_a = init_t.clone()
    c_b = torch.empty_like(c_a)
    if args.warmup > 0:
        kd.kmeans_lloyd_driver_ex(X_t, w, c_a, labels, c_b, 0.0, False, args.warmup,
                                  args.precision, stream)
    torch.cuda.synchronize()

    # timed: exactly --steps iterations from the seeded init, tol = 0
    c_a = init_t.clone()
    c_b = torch.empty_like(c_a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n_iter, inertia, stats = kd.kmeans_lloyd_driver_ex(
        X_t, w, c_a, labels, c_b, 0.0, False, args.steps, args.precision, stream)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    assert n_iter == args.steps, (n_iter, args.steps)

    # error against FP64 (untimed): the last update against an FP64 update of
    # the labels one iteration earlier, the inertia against FP64
    lab_prev = torch.empty_like(labels)
    c_p, c_q = init_t.clone(), torch.empty_like(init_t)
    kd.kmeans_lloyd_driver_ex(X_t, w, c_p, lab_prev, c_q, 0.0, False, args.steps - 1,
                              args.precision, stream)
    err = fp64_errors(X_t, w, lab_prev, c_b, labels, inertia)
    del lab_prev, c_p, c_q

    iters_per_s = n_iter / el
    ms_step = 1e3 * el / n_iter
    kern_ms = stats["fused_kernel_ms"] / max(1, stats["fused_launches"])
    esz = X_t.element_size()
    bytes_per_launch = n * d * esz + n * esz + n * 4      # X, weights, labels
    flops_per_launch = 2.0 * n * k * d + 2.0 * n * d
    hbm_gbs = bytes_per_launch / (kern_ms * 1e-3) / 1e9
    tflops = flops_per_launch / (kern_ms * 1e-3) / 1e12
    if args.dtype == "f64":
        compute_peak = FP64_PEAK_TFLOPS
    elif args.precision == "fp32":
        compute_peak = FP32_PEAK_TFLOPS
    else:
        compute_peak = BF16_PEAK_TFLOPS
    t_hbm = bytes_per_launch / (HBM_PEAK_GBS * 1e9)
    t_cmp = flops_per_launch / (compute_peak * 1e12)
    here = os.path.dirname(os.path.abspath(__file__))
    # HBM bytes per launch measured with rocprofv3 for this workload (only
    # when the shape is the one profiled)
    traffic = (traffic_from_profile(os.path.join(here, "profiles",
                                                 "fused_step_n10M_d32_k256_pmc.md"))
               if (n, d, k, args.precision, args.dtype) == (N, D, K, "fp32", "f32")
               else None)
    if t_cmp >= t_hbm:
        roof = {"bound": "mfma", "achieved": round(tflops, 3), "peak": compute_peak,
                "unit": "TFLOP/s", "frac": round(tflops / compute_peak, 4)}
    else:
        roof = {"bound": "hbm", "achieved": round(hbm_gbs, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(hbm_gbs / HBM_PEAK_GBS, 4)}
    roof["traffic"] = traffic
    roof["kernel"] = "fused assignment+update step"
    roof["kernel_ms"] = round(kern_ms, 4)
    roof["hbm_frac"] = round(hbm_gbs / HBM_PEAK_GBS, 4)
    roof["algorithmic_bytes"] = bytes_per_launch
    roof["algorithmic_flops"] = flops_per_launch

    cpu = None
    if not args.no_cpu_baseline and args.dtype == "f32":
        cpu = cpu_baseline(X_t, init_t, n, d, k, args.cpu_seconds)

    line = {
        "metric": METRIC,
        "value": round(iters_per_s, 3),
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": (args.dtype if args.precision == "fp32"
                  else f"{args.precision} cross term + f32 re-score"),
        "data": "synthetic",
        "config": {"workload": f"plain Lloyd, n={n:,} d={d} k={k} Gaussian blobs, "
                               f"tol=0, one iteration per step",
                   "n_samples": n, "n_features": d, "n_clusters": k,
                   "precision": args.precision},
        "roofline": roof,
        "cpu_baseline": cpu,
        "inertia": inertia,
        "error_vs_fp64": err,
        "path": stats.get("path"),
        "rescored_per_iter": stats.get("n_rescored", 0) / max(1, n_iter),
    }
    if not args.no_extras and args.precision == "fp32" and args.dtype == "f32":
        # beside the plain FP32 figure: the same fit with the BF16 cross term
        # (HBM roofline), the batched many-fit call and the served stream
        line["mixed_bf16"] = run_c1_mixed(args, X_t, w, init_t, labels, c_b, inertia)
        line["reference_api_fused_step"] = run_c1_fused_step(args, X_t, w, init_t)
        line["batched_c3"] = run_c3(args.fits, "bf16")
        line["served_c4"] = run_c4(args.extra_requests, args.rate, "bf16", "bf16")
        # arrivals faster than either form serves them: capacity, not the
        # arrival rate, bounds both fits/s figures
        line["served_c4_saturated"] = run_c4(args.extra_requests, args.saturated_rate,
                                             "bf16", "bf16")
    print(json.dumps(line), flush=True)


def run_c1_mixed(args, X_t, w, init_t, labels_fp32, cent_fp32, inertia_fp32):
    """The timed fit again with the BF16 cross term + certified FP32 re-score:
    iterations/s, the assignment+update kernels against the HBM roof, and
    bitwise equality with the FP32 fit."""
    import torch
    import kmeans_dpcpp_amd as kd
    stream = torch.cuda.current_stream()
    n, d = X_t.shape[1], X_t.shape[0]
    lab = torch.empty_like(labels_fp32)
    c_a, c_b = init_t.clone(), torch.empty_like(init_t)
    if args.warmup > 0:
        kd.kmeans_lloyd_driver_ex(X_t, w, c_a, lab, c_b, 0.0, False, args.warmup, "bf16",
                                  stream)
        c_a = init_t.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it, inertia, st = kd.kmeans_lloyd_driver_ex(X_t, w, c_a, lab, c_b, 0.0, False,
                                                args.steps, "bf16", stream)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    kern_ms = st["fused_kernel_ms"] / max(1, st["fused_launches"])
    bytes_per_launch = n * d * 4 + n * 4 + n * 4
    gbs = bytes_per_launch / (kern_ms * 1e-3) / 1e9
    return {
        "iters_per_s": round(it / el, 3),
        "ms_per_step": round(1e3 * el / it, 4),
        "path": st.get("path"),
        "rescored_per_iter": round(st.get("n_rescored", 0) / max(1, it), 1),
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "kernel": "BF16 assignment + update, then the FP32 re-score",
                     "kernel_ms": round(kern_ms, 4),
                     "algorithmic_bytes": bytes_per_launch},
        "labels_identical_to_fp32": bool(torch.equal(lab, labels_fp32)),
        "centroids_identical_to_fp32": bool(torch.equal(c_b.view(torch.int32),
                                                        cent_fp32.view(torch.int32))),
        "inertia_identical_to_fp32": inertia == inertia_fp32,
    }


def run_c1_fused_step(args, X_t, w, init_t):
    """The reference's own entry point on the same data: fused_lloyd_single_step
    (labels + private-copy sums from the seeded centroids), timed per call
    with HIP events on the stream it runs on; its labels are checked against
    the assignment entry point's."""
    import torch
    import kmeans_dpcpp_amd as kd
    stream = torch.cuda.current_stream()
    d, n = X_t.shape
    k = init_t.shape[1]
    hn = torch.empty(k, dtype=torch.float32, device=X_t.device)
    kd.half_l2_norm_squared(init_t, hn, 256, stream)
    n_copies = 4
    ids = torch.empty(n, dtype=torch.int32, device=X_t.device)
    cp = torch.zeros((n_copies, d, k), dtype=torch.float32, device=X_t.device)
    scp = torch.zeros((n_copies, k), dtype=torch.float32, device=X_t.device)
    calls = max(1, args.steps)
    for _ in range(2):
        kd.fused_lloyd_single_step(X_t, w, init_t, hn, ids, cp, scp, 8, 128, stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(calls):
        kd.fused_lloyd_single_step(X_t, w, init_t, hn, ids, cp, scp, 8, 128, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / calls
    ref = torch.empty_like(ids)
    kd.assignment(X_t, init_t, hn, ref, 8, 128, stream)
    torch.cuda.synchronize()
    counts = scp.double().sum(0) / (calls + 2)
    flops = 2.0 * n * k * d + 2.0 * n * d
    return {
        "ms_per_call": round(ms, 4),
        "n_copies": n_copies,
        "frac_fp32_peak": round(flops / (ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4),
        "labels_equal_assignment": bool(torch.equal(ids, ref)),
        "counts_equal_bincount": bool(torch.equal(
            counts, torch.bincount(ref.long(), minlength=k).double())),
    }


if __name__ == "__main__":
    main()
