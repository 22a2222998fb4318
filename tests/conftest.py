import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "kats.npz")


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def torch_dev():
    """ROCm device for GPU tests (fails loudly when no GPU is present)."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected but no ROCm device is visible")
    return torch.device("cuda", 0)


@pytest.fixture(scope="session")
def kd():
    import kmeans_dpcpp_amd
    return kmeans_dpcpp_amd


def to_dev(a, dtype=None, device="cuda"):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a if dtype is None else a.astype(dtype)))
    return t.to(device)


def to_np(t):
    return t.detach().cpu().numpy()


CORNERS = np.array([[1, 1, 1], [1, 1, -1], [1, -1, 1], [-1, 1, 1],
                    [1, -1, -1], [-1, 1, -1], [-1, -1, 1], [-1, -1, -1]],
                   dtype=np.float64)


def scal(t):
    """Python scalar of a one-element device array."""
    return to_np(t).reshape(-1)[0]


def check_label_mismatches(X_t, C_t, got, exp, limit):
    """Labels of a GPU path (`got`) against the C restatement's (`exp`, same
    FP32 centroids C_t): every sample where they differ must be a genuine
    FP32 near-tie.  For such a sample with labels a != b, the FP64 pseudo
    scores hn[c] - C[:, c].x of a and b may differ by at most the sum of both
    paths' FP32 rounding bounds, 2 (d + 3) 2^-24 (S_a + S_b) with
    S_c = sum_f |C[f, c] x_f| + hn[c]; else one of the paths mislabelled it.
    Returns the number of differing samples and asserts it is <= limit."""
    diff = np.flatnonzero(np.asarray(got) != np.asarray(exp))
    if diff.size:
        X = np.asarray(X_t[:, diff], dtype=np.float64)
        C = np.asarray(C_t, dtype=np.float64)
        d = X.shape[0]
        a, b = np.asarray(got)[diff].astype(np.int64), np.asarray(exp)[diff].astype(np.int64)
        hn = 0.5 * np.einsum("fc,fc->c", C, C)
        sa = hn[a] - np.einsum("fs,fs->s", C[:, a], X)
        sb = hn[b] - np.einsum("fs,fs->s", C[:, b], X)
        mag = (np.einsum("fs,fs->s", np.abs(C[:, a]), np.abs(X)) + hn[a]
               + np.einsum("fs,fs->s", np.abs(C[:, b]), np.abs(X)) + hn[b])
        bound = 2 * (d + 3) * 2.0 ** -24 * mag
        bad = np.abs(sa - sb) > bound
        assert not bad.any(), (int(bad.sum()), diff[bad][:5],
                               (np.abs(sa - sb) / bound)[bad][:5])
    assert diff.size <= limit, (diff.size, limit)
    return int(diff.size)


def fp64_update(X_t, w, labels, k):
    """FP64 centroids (d, k) and weights (k) of one Lloyd update from given
    labels (chunked over samples; empty clusters: NaN centroid)."""
    X_t = np.asarray(X_t)
    d, n = X_t.shape
    lab = np.asarray(labels).astype(np.int64)
    w = np.asarray(w, dtype=np.float64)
    cnt = np.bincount(lab, w, minlength=k)
    sums = np.empty((d, k))
    for f in range(d):
        sums[f] = np.bincount(lab, X_t[f].astype(np.float64) * w, minlength=k)
    with np.errstate(invalid="ignore", divide="ignore"):
        return sums / cnt, cnt


def fp64_inertia(X_t, w, C_t, labels, chunk=1 << 20):
    """sum_s w_s |x_s - C[:, l_s]|^2 in FP64, chunked over samples."""
    X_t = np.asarray(X_t)
    C = np.asarray(C_t, dtype=np.float64)
    lab = np.asarray(labels).astype(np.int64)
    w = np.asarray(w, dtype=np.float64)
    tot = 0.0
    for s in range(0, X_t.shape[1], chunk):
        e = min(X_t.shape[1], s + chunk)
        diff = X_t[:, s:e].astype(np.float64) - C[:, lab[s:e]]
        tot += float(np.dot(np.einsum("fs,fs->s", diff, diff), w[s:e]))
    return tot
