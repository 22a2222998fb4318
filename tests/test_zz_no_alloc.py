"""GPU test of kmd_fused_lloyd_single_step_noalloc, the fused step for callers
that cannot allocate (its working data lives in a 1 MiB device array per
GPU).  Called through the C ABI (ctypes) and compared with
kmd_fused_lloyd_single_step on the same inputs, two calls in a row, the
second smaller than the first, for the tiled shapes (centroid tiles in the
array), the LDS-resident shapes and a float64 shape (no working data);
shapes whose working data exceed the array are refused before any device
work.

Bar: labels and unit-weight counts bit-exact; centroid sums within 1e-6 of
their scale (the per-block flush into the copies is a float atomic add, so
its order, and the last bits, may differ between any two calls)."""
import ctypes

import numpy as np
import pytest

from oracle import lloyd_oracle as lo
from tests.conftest import to_dev, to_np

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SIG = [ctypes.c_int, ctypes.c_int] + [ctypes.c_int64] * 6 + [ctypes.c_void_p] * 8


@pytest.fixture(scope="module")
def lib():
    from kmeans_dpcpp_amd.build import library_path
    so = ctypes.CDLL(library_path())
    for name in ("kmd_fused_lloyd_single_step", "kmd_fused_lloyd_single_step_noalloc"):
        fn = getattr(so, name)
        fn.argtypes = SIG
        fn.restype = ctypes.c_int
    so.kmd_last_error.restype = ctypes.c_char_p
    return so


def call(lib, name, X, w, C, hn, labels, copies, sizes):
    dt = 0 if X.dtype == torch.float32 else 1
    it = 0 if labels.dtype == torch.int32 else 1
    d, n = X.shape
    k = C.shape[1]
    stream = torch.cuda.current_stream().cuda_stream
    rc = getattr(lib, name)(dt, it, n, d, k, copies.shape[0], 8, 128,
                            X.data_ptr(), w.data_ptr(), C.data_ptr(), hn.data_ptr(),
                            labels.data_ptr(), copies.data_ptr(), sizes.data_ptr(),
                            stream)
    assert rc == 0, lib.kmd_last_error().decode()


def one_case(lib, n, d, k, n_copies, dtype, seed):
    print(f"case n={n} d={d} k={k} copies={n_copies} dtype={dtype.__name__}", flush=True)
    Xt, init = lo.gaussian_blobs(n, d, k, seed=seed, dtype=dtype)
    X, C = to_dev(Xt), to_dev(init)
    w = torch.ones(n, dtype=X.dtype, device=X.device)
    hn = to_dev(lo.half_l2_norm_squared(init))
    out = {}
    for name in ("kmd_fused_lloyd_single_step", "kmd_fused_lloyd_single_step_noalloc"):
        labels = torch.full((n,), -1, dtype=torch.int32, device=X.device)
        copies = torch.full((n_copies, d, k), 0.5, dtype=X.dtype, device=X.device)
        sizes = torch.full((n_copies, k), 2.0, dtype=X.dtype, device=X.device)
        call(lib, name, X, w, C, hn, labels, copies, sizes)
        torch.cuda.synchronize()
        out[name] = (to_np(labels), to_np(copies).astype(np.float64).sum(axis=0),
                     to_np(sizes).astype(np.float64).sum(axis=0))
    (la, ca, sa), (lb, cb, sb) = out.values()
    np.testing.assert_array_equal(la, lb)
    np.testing.assert_array_equal(sa, sb)
    assert np.abs(ca - cb).max() <= 1e-6 * np.abs(ca).max()
    # and both are the update of their labels (accumulated into the copies)
    lab = la.astype(np.int64)
    np.testing.assert_array_equal(sb - 2.0 * n_copies, np.bincount(lab, minlength=k))
    exp = np.stack([np.bincount(lab, Xt[f].astype(np.float64), minlength=k)
                    for f in range(d)]) + 0.5 * n_copies
    assert np.abs(cb - exp).max() <= 1e-5 * np.abs(exp).max()


def test_noalloc_fused_step_equals_the_allocating_one(lib, torch_dev):
    # LDS-resident shapes (no working data): a large call, then a smaller one
    one_case(lib, 300_000, 32, 256, 8, np.float32, seed=1)
    one_case(lib, 50_001, 14, 127, 3, np.float32, seed=2)
    # tiled shapes (centroid tiles in the array): large, then smaller
    one_case(lib, 120_000, 128, 1024, 2, np.float32, seed=3)
    one_case(lib, 20_011, 70, 300, 1, np.float32, seed=4)
    # float64: the generic kernel, no working data
    one_case(lib, 30_000, 9, 40, 2, np.float64, seed=5)


def test_noalloc_refuses_shapes_beyond_its_array(lib, torch_dev):
    dev = torch.device("cuda", 0)
    d, k, n = 128, 4096, 64       # centroid tiles of 2.1 MB > 1 MiB
    X = torch.zeros((d, n), device=dev)
    C = torch.zeros((d, k), device=dev)
    labels = torch.zeros(n, dtype=torch.int32, device=dev)
    copies = torch.zeros((1, d, k), device=dev)
    sizes = torch.zeros((1, k), device=dev)
    w = torch.ones(n, device=dev)
    hn = torch.zeros(k, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.kmd_fused_lloyd_single_step_noalloc(
        0, 0, n, d, k, 1, 8, 128, X.data_ptr(), w.data_ptr(), C.data_ptr(), hn.data_ptr(),
        labels.data_ptr(), copies.data_ptr(), sizes.data_ptr(), stream)
    assert rc == 1
    assert "scratch" in lib.kmd_last_error().decode()
