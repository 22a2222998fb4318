"""Writes tests/golden/kats.npz: the known-answer cases of the reference's
own test-suite (tests/test_kernels.py), with seeded inputs.

Each entry is data: the inputs those tests build and the expected outputs
they assert, taken from the literal expectations and closed forms in
tests/test_kernels.py (line numbers in the keys' comments).  Where the
reference draws unseeded random inputs, a fixed seed is used here; the
expected values follow from the same closed forms.

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import os

import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kats.npz")

CORNERS = np.array([[1, 1, 1], [1, 1, -1], [1, -1, 1], [-1, 1, 1],
                    [1, -1, -1], [-1, 1, -1], [-1, -1, 1], [-1, -1, -1]],
                   dtype=np.float64)


def clouds(cloud_size, seed, dtype, rng_kind="legacy"):
    """8 Gaussian clouds (sigma 0.1) around the cube corners, rows = samples
    (tests/test_kernels.py:566-572 and the assignment/step/driver tests)."""
    if rng_kind == "default_rng":
        rs = np.random.default_rng(seed=seed)
        parts = [rs.normal(0, 0.1, size=(cloud_size, 3)).astype(dtype) + p.astype(dtype)
                 for p in CORNERS]
    else:
        rs = np.random.RandomState(seed)
        parts = [rs.normal(0, 0.1, size=(cloud_size, 3)).astype(dtype) + p.astype(dtype)
                 for p in CORNERS]
    return np.concatenate(parts, axis=0)


def main():
    g = {}
    # broadcast_divide -- tests/test_kernels.py:14-43
    g["bdiv_X"] = np.full((16, 5), 27.0)
    g["bdiv_y"] = np.array([1, 3, 9, 3, 1], dtype=np.float64)
    g["bdiv_expected"] = np.full((16, 5), np.array([27, 9, 3, 9, 27], dtype=np.float64))

    # half_l2_norm_squared -- :46-83
    X = np.empty((3, 6))
    X[:, :4] = 2.0
    X[:, 4:] = 1.0
    g["hn_X"] = X
    g["hn_expected"] = np.array([6., 6., 6., 6., 1.5, 1.5])

    # reduce_centroids_data, no empty cluster -- :86-129
    rs = np.random.RandomState(86)
    g["rcd_sizes_copies"] = np.array([[2, 1, 3, 1], [1, 2, 3, 1], [0, 1, 0, 1]],
                                     dtype=np.float64)
    g["rcd_cent_copies"] = rs.randn(3, 2, 4)
    g["rcd_canary"] = np.array(10)
    g["rcd_expected_n_empty"] = np.array(0)
    # reduce_centroids_data, one empty cluster (#3) -- :174-217
    g["rcde_sizes_copies"] = np.array([[2, 1, 3, 0], [1, 2, 3, 0], [1, 1, 0, 0]],
                                      dtype=np.float64)
    cc = rs.randn(3, 2, 4)
    np.put(cc, np.arange(3, 12, step=4), np.zeros(3))
    g["rcde_cent_copies"] = cc
    g["rcde_expected_n_empty"] = np.array(1)
    g["rcde_expected_list"] = np.array([3, 10, 10, 10])

    # compute_threshold -- :265-326 (n = 255*23)
    rs = np.random.RandomState(265)
    data = rs.randn(255 * 23)
    g["thr_data"] = data
    # expected values are computed per dtype from these closed forms:
    #   topk=1 -> max(data); topk=2 -> np.partition(data, n-2)[n-2]

    # select_samples_far_from_centroid -- :329-400 (n = 1e5, Gamma(4), 3 empties)
    rs = np.random.RandomState(329)
    g["sel_dist"] = rs.standard_gamma(4, size=10 ** 5).astype(np.float32)
    g["sel_n_empty"] = np.array(3)

    # relocate_empty_clusters -- :403-457
    g["rel_X"] = np.array([[+1, +1, +1], [+1, +1, -1], [+1, -1, +1], [-1, +1, +1],
                           [-1, -1, +1], [-1, +1, -1], [+1, -1, -1], [-1, -1, -1]],
                          dtype=np.float64)
    g["rel_C"] = np.array([[0.11, -0.1, 0.09], [5, 5, 5]])
    g["rel_empty_list"] = np.array([1, 0])
    g["rel_sizes"] = np.array([8, 0], dtype=np.float64)
    g["rel_expected_centroid_t"] = np.array([[0.11 + 1, -1], [-0.1 - 1, 1],
                                             [0.09 + 1, -1]])
    g["rel_expected_sizes"] = np.array([7, 1], dtype=np.float64)
    g["rel_moved_sample"] = np.array(5)

    # compute_centroid_shifts_squared -- :519-558
    g["shift_a"] = np.array([[1, -5], [2, -4], [3, -3], [4, -2], [5, -1]], dtype=np.float64)
    g["shift_b"] = np.array([[2, -4], [3, -3], [4, -2], [5, -1], [6, 1]], dtype=np.float64)
    g["shift_expected"] = np.array([5, 8], dtype=np.float64)

    # compute_centroid_to_sample_distances -- :561-642 (8 points per corner)
    g["dist_X"] = clouds(8, 561, np.float64)
    g["dist_C"] = CORNERS.copy()
    # assignment -- :645-777 (cloud 1734 in fp32, 32 in fp64)
    g["assign_X_1734"] = clouds(1734, 645, np.float32)
    g["assign_X_32"] = clouds(32, 735, np.float64)
    # compute_inertia -- :780-871 (cloud 32)
    g["inertia_X"] = clouds(32, 780, np.float64)
    # reduce_vector_blocking -- :874-893
    g["rv_data"] = np.arange(100, dtype=np.float64)
    g["rv_expected"] = np.array(4950.0)
    # fused_lloyd_single_step -- :896-1027 (cloud 16, 4 copies)
    g["step_X"] = clouds(16, 896, np.float64)
    g["step_n_copies"] = np.array(4)
    # kmeans_lloyd_driver -- :1030-1147 (default_rng(12345), cloud 32,
    # tol 1e-6, max_iter 255): labels = repeat(arange(8), 32), n_iter == 2
    g["drv_X_f32"] = clouds(32, 12345, np.float32, rng_kind="default_rng").astype(np.float64)
    g["drv_X_f64"] = clouds(32, 12345, np.float64, rng_kind="default_rng")
    g["drv_expected_n_iter"] = np.array(2)
    g["drv_tol"] = np.array(1e-6)
    g["drv_max_iter"] = np.array(255)

    np.savez_compressed(OUT, **g)
    print("wrote", OUT, "with", len(g), "arrays")


if __name__ == "__main__":
    main()
