"""Host-side tests of the drop-in boundary (no GPU work is launched):
the C-ABI library exports every function include/kmeans_dpcpp_amd.h
declares, the pybind11 module keeps the reference's 14 names and positional
parameter order, and argument validation raises before any device call."""
import ctypes
import os
import re

import numpy as np
import pytest

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "kmeans_dpcpp_amd.h")

# positional parameters of the reference module, python_api/_kmeans_lloyd.cpp:1277-1449
REFERENCE_SIGNATURES = {
    "broadcast_divide": ["divident", "divisor", "work_group_size", "sycl_queue", "depends"],
    "half_l2_norm_squared": ["centroids", "centroids_half_l2_norm_squared",
                             "work_group_size", "sycl_queue", "depends"],
    "reduce_centroids_data": ["cluster_sizes_private_copies", "centroids_t_private_copies",
                              "out_cluster_sizes", "out_centroids_t",
                              "out_empty_clusters_list", "out_n_empty_clusters",
                              "work_group_size", "sycl_queue", "depends"],
    "compute_threshold": ["data", "topk", "threshold", "sycl_queue", "depends"],
    "select_samples_far_from_centroid": ["n_selected", "distance_to_centroid", "threshold",
                                         "selected_samples_idx", "n_selected_gt_threshold",
                                         "n_selected_eq_threshold", "work_group_size",
                                         "sycl_queue", "depends"],
    "relocate_empty_clusters": ["n_empty_clusters", "X_t", "sample_weights", "assignment_id",
                                "empty_clusters_list", "sq_dist_to_nearest_centroid",
                                "centroid_t", "cluster_sizes", "per_sample_inertia",
                                "work_group_size", "sycl_queue", "depends"],
    "compute_centroid_shifts_squared": ["centroid_t", "new_centroid_t", "out_centroid_shifts",
                                        "sycl_queue", "depends"],
    "compute_centroid_to_sample_distances": ["X_t", "centroid_t", "euclidean_distances_t",
                                             "work_group_size", "centroids_window_height",
                                             "sycl_queue", "depends"],
    "assignment": ["X_t", "centroids_t", "centroids_half_l2_norm", "assignment_id",
                   "centroids_window_height", "work_group_size", "sycl_queue", "depends"],
    "compute_inertia": ["X_t", "sample_weight", "centroid_t", "assignment_id",
                        "out_per_sample_inertia", "work_group_size", "sycl_queue", "depends"],
    "reduce_vector_blocking": ["data", "sycl_queue", "depends"],
    "fused_lloyd_single_step": ["X_t", "sample_weight", "centroids_t", "centroids_half_l2_norm",
                                "assignments_idx", "new_centroids_t_private_copies",
                                "cluster_sizes_private_copies", "centroids_window_height",
                                "work_group_size", "sycl_queue", "depends"],
    "compute_number_of_private_copies": ["array", "n_samples", "n_features", "n_clusters",
                                         "centroids_private_copies_max_cache_occupancy",
                                         "work_group_size"],
    "kmeans_lloyd_driver": ["X_t", "sample_weight", "init_centroid_t", "assignments_id",
                            "res_centroids_t", "tol", "verbose", "max_iter",
                            "centroids_window_height", "work_group_size",
                            "centroids_private_copies_max_cache_occupancy", "sycl_queue",
                            "depends"],
}


@pytest.fixture(scope="module")
def lib():
    from kmeans_dpcpp_amd.build import library_path
    return ctypes.CDLL(library_path())


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(kmd_\w+)\s*\(", text,
                                 flags=re.M)))


def test_header_declares_the_reference_surface():
    names = declared_functions()
    for ref in REFERENCE_SIGNATURES:
        assert f"kmd_{ref}" in names, ref
    assert "kmd_iterative_merge_sort" in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_pybind_module_keeps_reference_names_and_order(kd):
    for name, params in REFERENCE_SIGNATURES.items():
        fn = getattr(kd, name)
        sig = fn.__doc__.splitlines()[0]
        got = re.findall(r"(\w+):", sig[sig.index("(") + 1:sig.rindex(")")])
        assert got == params, (name, got)
    assert set(REFERENCE_SIGNATURES) <= set(kd.__all__)


def _err(lib):
    lib.kmd_last_error.restype = ctypes.c_char_p
    return lib.kmd_last_error().decode()


def test_c_abi_rejects_bad_arguments_before_device_work(lib):
    i64, p = ctypes.c_int64, ctypes.c_void_p
    # unsupported dtype
    rc = lib.kmd_broadcast_divide(7, i64(2), i64(2), i64(32), p(0), p(0), p(0))
    assert rc == 1 and "Unsupported elemental data type" in _err(lib)
    # occupancy outside (0, 1)
    out = ctypes.c_int64(-1)
    rc = lib.kmd_compute_number_of_private_copies(0, 0, i64(10), i64(3), i64(4),
                                                  ctypes.c_double(1.5), i64(128),
                                                  ctypes.byref(out))
    assert rc == 1 and "strictly between 0 and 1" in _err(lib)
    # n_selected == 0
    rc = lib.kmd_select_samples_far_from_centroid(0, 0, i64(0), i64(10), i64(32), p(0),
                                                  p(0), p(0), p(0), p(0), p(0))
    assert rc == 1 and "must be positive" in _err(lib)
    # n_empty == 0
    rc = lib.kmd_relocate_empty_clusters(0, 0, i64(8), i64(3), i64(2), i64(0), i64(32),
                                         p(0), p(0), p(0), p(0), p(0), p(0), p(0), p(0),
                                         p(0))
    assert rc == 1 and "non-zero" in _err(lib)
    # negative tolerance
    n_it, inert = ctypes.c_int64(0), ctypes.c_double(0)
    rc = lib.kmd_kmeans_lloyd_driver(0, 0, i64(8), i64(3), i64(2), p(0), p(0), p(0), p(0),
                                     p(0), ctypes.c_double(-1.0), 0, i64(10), i64(8),
                                     i64(128), ctypes.c_double(0.5), ctypes.byref(n_it),
                                     ctypes.byref(inert), p(0), p(0), p(0))
    assert rc == 1 and "Tolerance must be non-negative" in _err(lib)
    # topk == 0 is a no-op (python_api/_kmeans_lloyd.cpp:312-316)
    assert lib.kmd_compute_threshold(0, i64(5), p(0), i64(0), p(0), p(0)) == 0


def test_pybind_rejects_host_arrays(kd):
    import torch
    X = torch.zeros(16, 5)
    with pytest.raises(ValueError, match="ROCm device array"):
        kd.broadcast_divide(X, torch.ones(5), 32, None)
    with pytest.raises(ValueError, match="DLPack"):
        kd.half_l2_norm_squared([1, 2], [3], 32, None)
    with pytest.raises(ValueError, match="n_selected"):
        kd.select_samples_far_from_centroid(0, X, X, X, X, X, 32, None)
    with pytest.raises(ValueError, match="sycl_queue"):
        kd.broadcast_divide(X, torch.ones(5), 32, "not a stream")


def test_default_event_is_complete(kd):
    ev = kd.Event()
    assert ev.is_complete
    ev.wait()


def test_c_abi_rejects_null_arrays(lib):
    """A NULL device array with a non-zero element count is KMD_EVALUE, not a
    device fault (the checks run before any HIP call)."""
    i64, p, dbl = ctypes.c_int64, ctypes.c_void_p, ctypes.c_double
    fake = p(0x1000)          # never dereferenced: a NULL elsewhere fails first
    n_it, inert = ctypes.c_int64(0), ctypes.c_double(0)
    calls = {
        "assignment": lambda: lib.kmd_assignment(0, 0, i64(8), i64(3), i64(2), i64(8),
                                                 i64(128), p(0), fake, fake, fake, p(0)),
        "compute_inertia": lambda: lib.kmd_compute_inertia(
            0, 0, i64(8), i64(3), i64(2), i64(128), fake, p(0), fake, p(0), fake, p(0)),
        "fused_lloyd_single_step": lambda: lib.kmd_fused_lloyd_single_step(
            0, 0, i64(8), i64(3), i64(2), i64(1), i64(8), i64(128), fake, p(0), fake,
            fake, fake, fake, fake, p(0)),
        "kmeans_lloyd_driver": lambda: lib.kmd_kmeans_lloyd_driver(
            0, 0, i64(8), i64(3), i64(2), fake, p(0), fake, fake, fake, dbl(0.0), 0,
            i64(10), i64(8), i64(128), dbl(0.5), ctypes.byref(n_it), ctypes.byref(inert),
            p(0), p(0), p(0)),
        "kmeans_lloyd_driver_ex": lambda: lib.kmd_kmeans_lloyd_driver_ex(
            0, 0, 0, i64(8), i64(3), i64(2), p(0), fake, fake, fake, fake, dbl(0.0), 0,
            i64(10), ctypes.byref(n_it), ctypes.byref(inert), p(0), p(0), p(0), p(0),
            p(0), p(0), p(0)),
        "relocate_empty_clusters": lambda: lib.kmd_relocate_empty_clusters(
            0, 0, i64(8), i64(3), i64(2), i64(1), i64(32), fake, fake, fake, fake, fake,
            p(0), fake, fake, p(0)),
        "reduce_centroids_data": lambda: lib.kmd_reduce_centroids_data(
            0, 0, i64(2), i64(3), i64(4), i64(32), fake, p(0), fake, fake, fake, fake,
            p(0)),
        "compute_centroid_shifts_squared": lambda: lib.kmd_compute_centroid_shifts_squared(
            0, i64(3), i64(4), fake, fake, p(0), p(0)),
        "compute_centroid_to_sample_distances":
            lambda: lib.kmd_compute_centroid_to_sample_distances(
                0, i64(8), i64(3), i64(2), i64(128), i64(8), fake, p(0), fake, p(0)),
        "reduce_vector_blocking": lambda: lib.kmd_reduce_vector_blocking(
            0, i64(8), p(0), ctypes.byref(inert), p(0)),
        "iterative_merge_sort": lambda: lib.kmd_iterative_merge_sort(
            0, i64(8), fake, p(0), p(0)),
    }
    for name, call in calls.items():
        rc = call()
        assert rc == 1, name
        assert "null pointer" in _err(lib), (name, _err(lib))
    # a NULL weight is the documented unweighted inertia (not an error), and
    # empty arrays may be NULL
    assert lib.kmd_compute_centroid_shifts_squared(0, i64(0), i64(0), p(0), p(0), p(0),
                                                   p(0)) == 0
