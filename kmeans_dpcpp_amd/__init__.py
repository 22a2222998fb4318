"""MI355X (gfx950) HIP implementation of the Lloyd k-means entry points of
kmeans_dpcpp (kmeans_dpcpp/__init__.py:1-33 re-exports the same 14 names).

Arrays are passed through DLPack (ROCm device arrays such as torch tensors on
a ROCm device); the reference's `sycl_queue` argument takes a stream-like
(None, an int hipStream_t, a torch.cuda.Stream or :class:`Stream`).

Importing this package loads the native module built by
``python kmeans_dpcpp_amd/build.py``; there is no fallback implementation.
"""
from __future__ import annotations

import os as _os

_HERE = _os.path.dirname(_os.path.abspath(__file__))

try:
    from . import _kmeans_dpcpp_amd as _native
except ImportError as _e:  # pragma: no cover - exercised only without a build
    raise ImportError(
        "kmeans_dpcpp_amd native module is not built; run "
        "`python kmeans_dpcpp_amd/build.py` (needs hipcc for gfx950): "
        f"{_e}") from _e

from ._kmeans_dpcpp_amd import (  # noqa: E402
    Engine,
    Event,
    assignment,
    autotune,
    broadcast_divide,
    column_mean_var,
    compute_centroid_shifts_squared,
    compute_centroid_to_sample_distances,
    compute_inertia,
    compute_number_of_private_copies,
    compute_threshold,
    fp8_roundtrip,
    fused_lloyd_single_step,
    gather_columns,
    half_l2_norm_squared,
    iterative_merge_sort,
    kmeans_lloyd_batched,
    kmeans_lloyd_driver,
    kmeans_lloyd_driver_ex,
    kmeans_plusplus_t,
    record_event,
    reduce_centroids_data,
    reduce_vector_blocking,
    relocate_empty_clusters,
    select_samples_far_from_centroid,
    set_tuning,
    transpose,
    version,
)
from .estimator import KMeans, kmeans_plusplus  # noqa: E402


class Stream:
    """A non-blocking HIP stream owned by the library.  Pass it (or its
    ``handle``) wherever the reference takes ``sycl_queue``."""

    def __init__(self):
        self.handle = _native._stream_create()

    def synchronize(self) -> None:
        _native._stream_synchronize(self.handle)

    def record_event(self) -> Event:
        return _native.record_event(self.handle)

    def close(self) -> None:
        if self.handle:
            _native._stream_destroy(self.handle)
            self.handle = 0

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass


def native_library_paths() -> list[str]:
    """Paths of the native shared objects this package loads."""
    from .build import library_path, pybind_module_path
    return [library_path(), pybind_module_path()]


__all__ = [
    "broadcast_divide",
    "half_l2_norm_squared",
    "reduce_centroids_data",
    "compute_threshold",
    "select_samples_far_from_centroid",
    "relocate_empty_clusters",
    "compute_centroid_shifts_squared",
    "compute_centroid_to_sample_distances",
    "assignment",
    "compute_inertia",
    "reduce_vector_blocking",
    "fused_lloyd_single_step",
    "compute_number_of_private_copies",
    "kmeans_lloyd_driver",
    # superset
    "iterative_merge_sort",
    "kmeans_lloyd_driver_ex",
    "Event",
    "Stream",
    "record_event",
    "version",
]
