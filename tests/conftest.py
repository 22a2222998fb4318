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
