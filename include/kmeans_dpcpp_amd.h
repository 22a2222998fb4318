/*
 * kmeans_dpcpp_amd.h -- C ABI of the MI355X (gfx950) Lloyd k-means library.
 *
 * Every entry point below replaces one function exported by the reference's
 * pybind11 module `_kmeans_dpcpp` (python_api/_kmeans_lloyd.cpp:1277-1449).
 * The reference passes dpctl usm_ndarray objects and a sycl::queue; this ABI
 * passes plain device pointers, element counts and an opaque HIP stream.
 *
 * Conventions
 *   - Layouts are the reference's feature-major ("transposed") layouts:
 *       X_t            (n_features, n_samples)          row-major
 *       centroids_t    (n_features, n_clusters)         row-major
 *       private copies (n_copies, n_features, n_clusters)
 *       distances      (n_clusters, n_samples)
 *   - `dtype` selects the floating type of data arrays (KMD_F32 / KMD_F64),
 *     `itype` the integer type of labels / index arrays (KMD_I32 / KMD_I64).
 *   - `stream` is a hipStream_t passed as void*; NULL is the null stream.
 *     Asynchronous entry points only enqueue work on `stream`.
 *   - Return value: KMD_OK, KMD_EVALUE (bad argument; the reference raises
 *     ValueError) or KMD_ERUNTIME (HIP failure).  kmd_last_error() returns a
 *     message for the last failure on the calling thread.
 *   - Every pointer must be device memory on the current HIP device, except
 *     where a parameter is documented as host memory.
 */
#ifndef KMEANS_DPCPP_AMD_H
#define KMEANS_DPCPP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { KMD_OK = 0, KMD_EVALUE = 1, KMD_ERUNTIME = 2 };
enum { KMD_F32 = 0, KMD_F64 = 1 };
enum { KMD_I32 = 0, KMD_I64 = 1 };

/* Precision of the sample-centroid cross term in the Lloyd kernels.
 * KMD_PREC_FP32    : exact FP32 products (vector FMA or FP32 MFMA).
 * KMD_PREC_BF16X3  : three-term BF16 split on MFMA, FP32 accumulate, then an
 *                    exact FP32 rescore of every sample whose best/second-best
 *                    margin is below the split's error bound.
 * KMD_PREC_BF16    : one-term BF16 on MFMA + the same bounded FP32 rescore.
 * KMD_PREC_FP8     : FP8 e4m3 on MFMA with a power-of-two scale per
 *                    32-centroid tile (and one for the samples), FP32
 *                    accumulate, + the bounded FP32 rescore (e4m3's error
 *                    bound; float32 data with n_features <= 128 on the
 *                    tiled kernels, else the FP32 form).
 * Labels are those of the FP32 form in every mode. */
enum { KMD_PREC_FP32 = 0, KMD_PREC_BF16X3 = 1, KMD_PREC_BF16 = 2, KMD_PREC_FP8 = 3 };

typedef void *kmd_stream_t;
typedef void *kmd_event_t;
typedef void (*kmd_print_fn)(const char *message, void *user_data);

/* ---- diagnostics and runtime helpers (no reference counterpart) -------- */
const char *kmd_last_error(void);
const char *kmd_version(void);
int kmd_device_count(int *count);
int kmd_set_device(int device);
int kmd_event_create(kmd_event_t *event);
int kmd_event_record(kmd_event_t event, kmd_stream_t stream);
int kmd_event_synchronize(kmd_event_t event);
int kmd_event_query(kmd_event_t event, int *done);
int kmd_event_elapsed_ms(kmd_event_t start, kmd_event_t stop, float *ms);
int kmd_event_destroy(kmd_event_t event);
int kmd_stream_wait_event(kmd_stream_t stream, kmd_event_t event);
int kmd_stream_synchronize(kmd_stream_t stream);
int kmd_stream_create(kmd_stream_t *stream);
int kmd_stream_destroy(kmd_stream_t stream);
/* out[i] = in[i] rounded to FP8 e4m3 and back, with the conversion the
 * KMD_PREC_FP8 kernels use (device arrays of n floats): lets a test pin the
 * rounding their error bound assumes. */
int kmd_fp8_roundtrip(int64_t n, const float *in, float *out, kmd_stream_t stream);

/* ---- the 14 reference entry points ------------------------------------ */

/* divident[f, c] /= divisor[c].  Replaces py_broadcast_divide
 * (python_api/_kmeans_lloyd.cpp:60-105) -> broadcast_division_kernel
 * (src/util_kernels.hpp:9-41).  n_rows = divident.shape[0],
 * n_cols = divident.shape[1] = divisor.shape[0]. */
int kmd_broadcast_divide(int dtype, int64_t n_rows, int64_t n_cols,
                         int64_t work_group_size, void *divident,
                         const void *divisor, kmd_stream_t stream);

/* out[c] = (sum_f centroids[f, c]^2) / 2, f ascending.  Replaces
 * py_half_l2_norm_squared (python_api/_kmeans_lloyd.cpp:108-151) ->
 * half_l2_norm_kernel (src/util_kernels.hpp:47-82). */
int kmd_half_l2_norm_squared(int dtype, int64_t n_features, int64_t n_clusters,
                             int64_t work_group_size, const void *centroids,
                             void *out, kmd_stream_t stream);

/* Sums private copies (copy index ascending) into out_centroids_t and
 * out_cluster_sizes; appends each cluster whose summed size is 0 to
 * out_empty_clusters_list at position out_n_empty_clusters[0]++ (ascending
 * cluster order; slots not written keep their values; writes past
 * list_len are dropped but still counted).  Replaces py_reduce_centroids_data
 * (python_api/_kmeans_lloyd.cpp:154-282) -> reduce_centroid_data_kernel
 * (src/util_kernels.hpp:84-161). */
int kmd_reduce_centroids_data(int dtype, int itype, int64_t n_copies,
                              int64_t n_features, int64_t n_clusters,
                              int64_t work_group_size,
                              const void *cluster_sizes_private_copies,
                              const void *centroids_t_private_copies,
                              void *out_cluster_sizes, void *out_centroids_t,
                              void *out_empty_clusters_list,
                              void *out_n_empty_clusters, kmd_stream_t stream);

/* threshold[0] = topk-th largest element of data (topk == 1: the maximum;
 * topk >= 2: ascending_sort(data)[n - topk]; topk == 0: nothing is done).
 * Replaces py_compute_threshold (python_api/_kmeans_lloyd.cpp:284-337) ->
 * compute_threshold_kernel (src/util_kernels.hpp:163-206). */
int kmd_compute_threshold(int dtype, int64_t n, const void *data, int64_t topk,
                          void *threshold, kmd_stream_t stream);

/* Indices with distance > threshold go to the front of selected_samples_idx
 * (ascending index order) and n_selected_gt_threshold[0] grows by their
 * count; indices with distance == threshold go to the back
 * (selected[n - n_eq_old], n_eq grows by one per entry, ascending index
 * order) until n_selected entries are selected in total.  Replaces
 * py_select_samples_far_from_centroid (python_api/_kmeans_lloyd.cpp:339-441)
 * -> select_samples_far_from_centroid_kernel (src/util_kernels.hpp:208-297). */
int kmd_select_samples_far_from_centroid(
    int dtype, int itype, int64_t n_selected, int64_t n_samples,
    int64_t work_group_size, const void *distance_to_centroid,
    const void *threshold, void *selected_samples_idx,
    void *n_selected_gt_threshold, void *n_selected_eq_threshold,
    kmd_stream_t stream);

/* For the j-th entry of empty_clusters_list (j < n_empty_clusters) the j-th
 * farthest sample s (sq_dist descending, index ascending) moves from its
 * cluster p = assignment_id[s]: centroid_t[:, p] -= w_s x_s,
 * centroid_t[:, e_j] = w_s x_s, cluster_sizes[p] -= w_s,
 * cluster_sizes[e_j] = w_s, per_sample_inertia[s] = 0 (centroid_t holds
 * unnormalised sums).  Replaces py_relocate_empty_clusters
 * (python_api/_kmeans_lloyd.cpp:443-589) -> relocate_empty_clusters
 * (src/util_kernels.hpp:392-476). */
int kmd_relocate_empty_clusters(
    int dtype, int itype, int64_t n_samples, int64_t n_features,
    int64_t n_clusters, int64_t n_empty_clusters, int64_t work_group_size,
    const void *X_t, const void *sample_weights, const void *assignment_id,
    const void *empty_clusters_list, const void *sq_dist_to_nearest_centroid,
    void *centroid_t, void *cluster_sizes, void *per_sample_inertia,
    kmd_stream_t stream);

/* out[c] = sum_f (old[f, c] - new[f, c])^2.  Replaces
 * py_compute_centroid_shifts_squared_kernel
 * (python_api/_kmeans_lloyd.cpp:591-664) ->
 * compute_centroid_shifts_squared_kernel (src/util_kernels.hpp:478-516). */
int kmd_compute_centroid_shifts_squared(int dtype, int64_t n_features,
                                        int64_t n_clusters,
                                        const void *old_centroid_t,
                                        const void *new_centroid_t,
                                        void *out_centroid_shifts,
                                        kmd_stream_t stream);

/* D[c, s] = sqrt(sum_f (C[f, c] - X[f, s])^2) for every shape (the
 * reference's window loops are only correct when
 * ceil(d/h) == ceil(k/32)).  Replaces py_compute_distances
 * (python_api/_kmeans_lloyd.cpp:666-745) -> compute_distances
 * (src/compute_euclidean_distance.hpp:12-113). */
int kmd_compute_centroid_to_sample_distances(
    int dtype, int64_t n_samples, int64_t n_features, int64_t n_clusters,
    int64_t work_group_size, int64_t centroids_window_height, const void *X_t,
    const void *centroid_t, void *euclidean_distances_t, kmd_stream_t stream);

/* labels[s] = argmin_c (hn[c] - sum_f C[f, c] X[f, s]), lowest c on ties,
 * label 0 when no value is below +inf.  Replaces py_assignment
 * (python_api/_kmeans_lloyd.cpp:747-839) -> assignment
 * (src/assignment.hpp:11-146). */
int kmd_assignment(int dtype, int itype, int64_t n_samples, int64_t n_features,
                   int64_t n_clusters, int64_t centroids_window_height,
                   int64_t work_group_size, const void *X_t,
                   const void *centroids_t, const void *centroids_half_l2_norm,
                   void *assignment_id, kmd_stream_t stream);

/* out[s] = w[s] * sum_f (X[f, s] - C[f, labels[s]])^2.  Replaces
 * py_compute_inertia (python_api/_kmeans_lloyd.cpp:841-939) ->
 * compute_inertia_kernel (src/compute_inertia.hpp:11-53).  A NULL
 * sample_weight gives the unweighted form (compute_uniform_weight_inertia_kernel,
 * src/compute_inertia.hpp:58-99). */
int kmd_compute_inertia(int dtype, int itype, int64_t n_samples,
                        int64_t n_features, int64_t n_clusters,
                        int64_t work_group_size, const void *X_t,
                        const void *sample_weight, const void *centroid_t,
                        const void *assignment_id, void *out_per_sample_inertia,
                        kmd_stream_t stream);

/* *result (host memory) = sum of data[0:n]; blocks until done.  Replaces
 * py_reduce_vector_blocking (python_api/_kmeans_lloyd.cpp:941-976) ->
 * reduce_vector_kernel_blocking (src/compute_inertia.hpp:104-135). */
int kmd_reduce_vector_blocking(int dtype, int64_t n, const void *data,
                               double *result, kmd_stream_t stream);

/* One fused Lloyd step: labels as kmd_assignment, then ACCUMULATES
 * w_s x_s into new_centroids_t_private_copies[copy][:, label] and w_s into
 * cluster_sizes_private_copies[copy][label] (callers zero them first; the
 * copy a sample lands in is unspecified; only the sum over copies is
 * defined).  Replaces py_fused_lloyd_single_step
 * (python_api/_kmeans_lloyd.cpp:980-1109) -> lloyd_single_step
 * (src/lloyd_single_step.hpp:37-209). */
int kmd_fused_lloyd_single_step(
    int dtype, int itype, int64_t n_samples, int64_t n_features,
    int64_t n_clusters, int64_t n_copies, int64_t centroids_window_height,
    int64_t work_group_size, const void *X_t, const void *sample_weight,
    const void *centroids_t, const void *centroids_half_l2_norm,
    void *assignments_idx, void *new_centroids_t_private_copies,
    void *cluster_sizes_private_copies, kmd_stream_t stream);

/* kmd_fused_lloyd_single_step without any device allocation.  No
 * counterpart in the reference: for callers that cannot allocate -- code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  Same arguments, kernels and results as
 * kmd_fused_lloyd_single_step; its working data (the prepared centroid
 * tiles of the shapes beyond LDS; the LDS-resident shapes need none) lives
 * in one 1 MiB device array per GPU, held until the
 * process ends.  Shapes that need more than 1 MiB (float32 with d <= 128 and
 * a centroid table beyond LDS: about 4 k (d + 1) bytes, e.g. d = 128 with
 * k > 2016) return KMD_EVALUE.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams or threads) race. */
int kmd_fused_lloyd_single_step_noalloc(
    int dtype, int itype, int64_t n_samples, int64_t n_features,
    int64_t n_clusters, int64_t n_copies, int64_t centroids_window_height,
    int64_t work_group_size, const void *X_t, const void *sample_weight,
    const void *centroids_t, const void *centroids_half_l2_norm,
    void *assignments_idx, void *new_centroids_t_private_copies,
    void *cluster_sizes_private_copies, kmd_stream_t stream);

/* *out = max(1, min(floor(cache * occupancy / (sizeof(T) k (d + 1))),
 * ceil(n / wg) * wg / 8)) with cache = L2 size of `device`.  Replaces
 * py_compute_number_of_private_copies (python_api/_kmeans_lloyd.cpp:1111-1147)
 * -> compute_number_of_private_copies (src/lloyd_single_step.hpp:12-35). */
int kmd_compute_number_of_private_copies(int dtype, int device,
                                         int64_t n_samples, int64_t n_features,
                                         int64_t n_clusters, double occupancy,
                                         int64_t work_group_size, int64_t *out);

/* Full Lloyd fit with the reference driver's semantics
 * (src/kmeans_lloyd_driver.hpp:21-431): iterate while
 * it < max_iter and sum(shift) > (T)tol; relocate empty clusters; final
 * labels and weighted inertia of the returned centroids.  init_centroids_t
 * is overwritten.  *n_iters and *total_inertia are host memory.  Blocks.
 * print_fn (may be NULL) receives the per-iteration line when verbose.
 * Replaces py_kmeans_lloyd_driver (python_api/_kmeans_lloyd.cpp:1149-1275). */
int kmd_kmeans_lloyd_driver(
    int dtype, int itype, int64_t n_samples, int64_t n_features,
    int64_t n_clusters, const void *X_t, const void *sample_weight,
    void *init_centroids_t, void *assignments_id, void *res_centroids_t,
    double tol, int verbose, int64_t max_iter, int64_t centroids_window_height,
    int64_t work_group_size, double occupancy, int64_t *n_iters,
    double *total_inertia, kmd_print_fn print_fn, void *print_user_data,
    kmd_stream_t stream);

/* ---- superset exports -------------------------------------------------- */

/* Stable ascending sort of data[0:n] into output (NaN sorts last).
 * Replaces iterative_merge_sort (src/iterative_merge_sort.hpp:204-368),
 * which the reference reaches only through compute_threshold. */
int kmd_iterative_merge_sort(int dtype, int64_t n, const void *input,
                             void *output, kmd_stream_t stream);

/* kmd_kmeans_lloyd_driver with a selectable cross-term precision
 * (KMD_PREC_*) and instrumentation.  No reference counterpart: the reference
 * has FP32/FP64 only.  In the BF16 modes (dtype must be KMD_F32) centroid
 * sums and the inertia reduction run in FP64.  Optional outputs (host
 * memory, may be NULL): *n_rescored = samples re-scored in FP32 summed over
 * iterations; *fused_kernel_ms = summed HIP-event duration of the fused-step
 * kernel launches (in the mixed modes: BF16 kernel + FP32 re-score),
 * *fused_launches = their number; *fused_path = 0 generic kernel, 1 FP32
 * MFMA, 2 BF16 + re-score, 3 BF16x3 + re-score, 4/5/6 the same three on the
 * tiled path, 7 FP8 + re-score on the tiled path (d <= 128 with centroids or sums beyond LDS: centroid tiles
 * streamed through LDS, update as a separate exact pass); a mixed mode falls
 * back to FP32 for shapes it does not serve or non-finite inputs. */
int kmd_kmeans_lloyd_driver_ex(
    int dtype, int itype, int precision, int64_t n_samples,
    int64_t n_features, int64_t n_clusters, const void *X_t,
    const void *sample_weight, void *init_centroids_t, void *assignments_id,
    void *res_centroids_t, double tol, int verbose, int64_t max_iter,
    int64_t *n_iters, double *total_inertia, int64_t *n_rescored,
    double *fused_kernel_ms, int64_t *fused_launches, int *fused_path,
    kmd_print_fn print_fn, void *print_user_data, kmd_stream_t stream);

/* One fit of a batch (float32 data, int32 labels; all pointers device
 * memory on the current device, as for kmd_kmeans_lloyd_driver). */
typedef struct kmd_fit {
  int64_t n_samples, n_features, n_clusters;
  const void *X_t;            /* (n_features, n_samples) */
  const void *sample_weight;  /* (n_samples) */
  void *init_centroids_t;     /* (n_features, n_clusters), overwritten */
  void *assignments_id;       /* (n_samples) int32 */
  void *res_centroids_t;      /* (n_features, n_clusters) */
  double tol;
  int64_t max_iter;
  int64_t n_iters;            /* out */
  double inertia;             /* out */
} kmd_fit;

/* Many independent Lloyd fits at once (no counterpart in the reference,
 * which fits one problem per kmeans_lloyd_driver call).  Each fit gets
 * exactly the results kmd_kmeans_lloyd_driver(KMD_F32, KMD_I32, ...) gives
 * it alone: labels, centroids, iteration count and inertia.  Fits with
 * d <= 64, k <= 256, k <= n and finite inputs run together: one persistent
 * launch per step of an iteration over a work list of (fit, tile) items,
 * per-fit convergence masks, the iteration captured once in a hipGraph and
 * replayed, `graph_replays` iterations between host checks of the masks
 * (<= 0: 4).  Other fits, and fits whose data range needs floating-point
 * centroid sums, run through the single-fit driver.  precision:
 * KMD_PREC_FP32, or KMD_PREC_BF16 (the batched assignment takes the cross
 * term in BF16 on the matrix core and re-scores in FP32 every sample whose
 * margin is below the error bound: the same labels, so the same fits, bit
 * for bit).  The first two iterations try the assignment kernels' two
 * launch shapes (the default threads per block, then half of them), timing
 * them, and the rest run with the faster; results never depend on it.
 * *graph_launches (may be NULL) = iterations issued (replays and trials),
 * *assign_shape (may be NULL) = the shape kept.  Blocking. */
int kmd_kmeans_lloyd_batched(int64_t n_fits, kmd_fit *fits, int precision,
                             int graph_replays, int64_t *graph_launches,
                             int *assign_shape, kmd_stream_t stream);

/* Continuous batching (no counterpart in the reference; for a serving loop
 * that admits fit requests while others run).  An engine holds a running
 * set of fits on one GPU.  kmd_engine_add joins a fit (the shapes of
 * kmd_kmeans_lloyd_batched's batched kernels; its arrays and the kmd_fit
 * struct must stay valid until the fit is reported finished);
 * kmd_engine_step advances every active fit by `iterations` Lloyd
 * iterations (use_graph: replay a captured hipGraph while the set is
 * unchanged) and finishes the fits that stopped, writing n_iters and
 * inertia into their kmd_fit; kmd_engine_take_finished returns their ids
 * (up to `capacity`; the rest stay pending).  Results equal
 * kmd_kmeans_lloyd_driver's for each fit alone, whatever else is in the
 * set.  kmd_engine_footprint = device bytes the engine holds per fit. */
typedef struct kmd_engine *kmd_engine_t;
int kmd_engine_create(kmd_engine_t *engine, kmd_stream_t stream);
int kmd_engine_destroy(kmd_engine_t engine);
int kmd_engine_footprint(int64_t n_samples, int64_t n_features,
                         int64_t n_clusters, int64_t *bytes);
int kmd_engine_add(kmd_engine_t engine, kmd_fit *fit, int64_t *id);
int kmd_engine_step(kmd_engine_t engine, int iterations, int use_graph,
                    int64_t *n_active);
int kmd_engine_take_finished(kmd_engine_t engine, int64_t *ids,
                             int64_t capacity, int64_t *count);
int kmd_engine_resident_bytes(kmd_engine_t engine, int64_t *bytes);
/* Assignment kernels of the engine's iterations: KMD_PREC_FP32 (default) or
 * KMD_PREC_BF16 (BF16 cross term + certified FP32 re-score; results equal
 * the FP32 engine's bit for bit).  Applies from the next step. */
int kmd_engine_set_precision(kmd_engine_t engine, int precision);

/* Launch autotuner (no counterpart in the reference, whose launch shape is
 * the caller's work_group_size / centroids_window_height).  Times the
 * driver's candidate plans for this problem -- LDS-resident or tiled
 * centroids (where each serves the shape) x, for the LDS-resident FP32
 * step, its launch shape (shape 0: 16 waves per block x one 32-sample
 * column block per wave; 1: 12 waves x two; 2: 8 waves x two) x persistent
 * blocks per CU {occupancy limit, 1, 2, 4} -- for `iters` Lloyd iterations
 * from init_centroids_t (not modified) and records the fastest for
 * (n_features, n_clusters, precision, log2 n_samples); later driver calls
 * of that shape use it.  Results never depend on the plan.  Writes the best
 * plan and up to `capacity` candidates (path: 1 LDS-resident, 2 tiled,
 * 0 generic; blocks_per_cu 0 = occupancy limit; shape; fused-step
 * ms/launch). */
int kmd_autotune(int dtype, int itype, int precision, int64_t n_samples,
                 int64_t n_features, int64_t n_clusters, const void *X_t,
                 const void *sample_weight, const void *init_centroids_t,
                 int iters, int *best_path, int *best_blocks_per_cu,
                 int *best_shape, double *best_ms, int *cand_path,
                 int *cand_blocks_per_cu, int *cand_shape, double *cand_ms,
                 int64_t capacity, int64_t *n_candidates, kmd_stream_t stream);
/* Process-wide override of the plan driver calls use (path -1 clears it:
 * back to the autotuner's table, else the automatic choice; shape as in
 * kmd_autotune). */
int kmd_set_tuning(int path, int blocks_per_cu, int shape);

/* ---- estimator front-end: data preparation and seeding ----------------
 * No counterpart in the reference, whose driver takes init_centroids_t from
 * its caller (python_api/_kmeans_lloyd.cpp:1437); these serve the caller the
 * reference's README (README.md:3-4) names: an sklearn-style estimator that
 * holds sample-major data. */

/* Per-feature mean and variance (FP64, two passes; variance = mean squared
 * deviation, numpy's np.var) of a row-major X (n_samples, n_features).
 * mean and var: device, n_features doubles; var may be NULL. */
int kmd_column_mean_var(int dtype, int64_t n_samples, int64_t n_features,
                        const void *X, double *mean, double *var,
                        kmd_stream_t stream);

/* dst (cols, rows) = src (rows, cols)^T - shift[col]; shift: device doubles
 * (cols) rounded to dtype before the subtraction, or NULL for a plain
 * transpose.  Turns sample-major X into X_t (and distances (k, n) into
 * (n, k)). */
int kmd_transpose(int dtype, int64_t rows, int64_t cols, const void *src,
                  const double *shift, void *dst, kmd_stream_t stream);

/* out_t[:, j] = X_t[:, indices[j]] for j < n_indices (indices: device
 * int64, each in [0, n_samples)). */
int kmd_gather_columns(int dtype, int64_t n_samples, int64_t n_features,
                       int64_t n_indices, const void *X_t,
                       const int64_t *indices, void *out_t,
                       kmd_stream_t stream);

/* Greedy k-means++ seeding, the algorithm of scikit-learn 1.7.2
 * (sklearn/cluster/_kmeans.py:174-277) with its random draws supplied:
 * uniforms (HOST memory, 1 + (n_clusters - 1) * n_local_trials doubles in
 * [0, 1), copied before the call returns): uniforms[0] picks the first
 * center (probability proportional to sample_weight, numpy choice's rule),
 * uniforms[1 + (c - 1) n_local_trials + j] the j-th candidate of center c
 * (probability proportional to w D^2); the candidate with the smallest
 * potential wins.  n_local_trials in [1, 16] (sklearn's default is
 * 2 + floor(ln n_clusters)).  Writes centers_t (n_features, n_clusters) and
 * the chosen sample indices (device int64, n_clusters).  Distances,
 * potentials and prefix sums are FP64. */
int kmd_kmeans_plusplus(int dtype, int64_t n_samples, int64_t n_features,
                        int64_t n_clusters, int64_t n_local_trials,
                        const void *X_t, const void *sample_weight,
                        const double *uniforms, void *centers_t,
                        int64_t *indices, kmd_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* KMEANS_DPCPP_AMD_H */
