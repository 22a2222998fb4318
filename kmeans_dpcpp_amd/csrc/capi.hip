// extern "C" entry points of libkmeans_dpcpp_amd.so (include/kmeans_dpcpp_amd.h).
// Each one validates its scalar arguments, dispatches on (dtype, itype) and
// enqueues the kernels on the caller's stream.  Exceptions never cross the C
// boundary: ValueError -> KMD_EVALUE, anything else -> KMD_ERUNTIME, with the
// message kept in a thread-local buffer for kmd_last_error().
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "lloyd_fast.hpp"

namespace kmd {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }

int device_cu_count() {
  static std::mutex mu;
  static std::vector<int> cache;
  int dev = 0;
  KMD_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if ((int)cache.size() <= dev) cache.resize(dev + 1, 0);
  if (cache[dev] == 0) {
    int cus = 0;
    KMD_HIP_CHECK(
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    cache[dev] = cus > 0 ? cus : 1;
  }
  return cache[dev];
}

void allow_full_lds(const void *kernel) {
  static std::mutex mu;
  static std::vector<const void *> done;
  std::lock_guard<std::mutex> lock(mu);
  for (const void *f : done)
    if (f == kernel) return;
  KMD_HIP_CHECK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    160 * 1024));
  done.push_back(kernel);
}

int64_t device_l2_bytes(int device) {
  int bytes = 0;
  KMD_HIP_CHECK(
      hipDeviceGetAttribute(&bytes, hipDeviceAttributeL2CacheSize, device));
  return bytes;
}

template <typename T, typename I>
void driver_lloyd(int64_t n, int64_t d, int64_t k, const T *X, const T *w,
                  T *init, I *labels, T *res, double tol, bool verbose,
                  int64_t max_iter, int64_t *n_iters_out, double *inertia_out,
                  kmd_print_fn print_fn, void *print_user, hipStream_t s,
                  int precision, DriverStats *stats);

}  // namespace kmd

using namespace kmd;

namespace {

template <typename F>
int guarded(F &&f) {
  try {
    f();
    return KMD_OK;
  } catch (const ValueError &e) {
    set_last_error(e.what());
    return KMD_EVALUE;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return KMD_ERUNTIME;
  } catch (...) {
    set_last_error("unknown error");
    return KMD_ERUNTIME;
  }
}

inline hipStream_t as_stream(kmd_stream_t s) {
  return reinterpret_cast<hipStream_t>(s);
}

inline void check_dtype(int dtype) {
  require(dtype == KMD_F32 || dtype == KMD_F64,
          "Unsupported elemental data type. Expecting single or double "
          "precision floating point numbers");
}
inline void check_types(int dtype, int itype) {
  require((dtype == KMD_F32 || dtype == KMD_F64) &&
              (itype == KMD_I32 || itype == KMD_I64),
          "Unsupported data types");
}
inline void check_ptr(const void *p, int64_t count, const char *name) {
  if (count > 0 && p == nullptr)
    throw ValueError(std::string("null pointer for non-empty array ") + name);
}
// The five arrays of one Lloyd fit (driver, batched fits, engine, autotuner).
inline void check_fit_arrays(int64_t n, int64_t d, int64_t k, const void *X_t,
                             const void *w, const void *init,
                             const void *labels, const void *res) {
  check_ptr(X_t, n * d, "X_t");
  check_ptr(w, n, "sample_weight");
  check_ptr(init, d * k, "init_centroids_t");
  check_ptr(labels, n, "assignments_id");
  check_ptr(res, d * k, "res_centroids_t");
}

// Calls fn.template operator()<T>() for the runtime dtype.
template <typename F>
void dispatch_t(int dtype, F &&fn) {
  if (dtype == KMD_F32)
    fn((float *)nullptr);
  else
    fn((double *)nullptr);
}
template <typename F>
void dispatch_ti(int dtype, int itype, F &&fn) {
  if (dtype == KMD_F32 && itype == KMD_I32)
    fn((float *)nullptr, (int32_t *)nullptr);
  else if (dtype == KMD_F32 && itype == KMD_I64)
    fn((float *)nullptr, (int64_t *)nullptr);
  else if (dtype == KMD_F64 && itype == KMD_I32)
    fn((double *)nullptr, (int32_t *)nullptr);
  else
    fn((double *)nullptr, (int64_t *)nullptr);
}

}  // namespace

extern "C" {

const char *kmd_last_error(void) { return g_last_error.c_str(); }

const char *kmd_version(void) { return "kmeans_dpcpp_amd 0.1.0 (gfx950)"; }

int kmd_device_count(int *count) {
  return guarded([&] {
    require(count != nullptr, "count must not be null");
    KMD_HIP_CHECK(hipGetDeviceCount(count));
  });
}

int kmd_set_device(int device) {
  return guarded([&] { KMD_HIP_CHECK(hipSetDevice(device)); });
}

int kmd_event_create(kmd_event_t *event) {
  return guarded([&] {
    require(event != nullptr, "event must not be null");
    hipEvent_t e;
    KMD_HIP_CHECK(hipEventCreate(&e));
    *event = reinterpret_cast<kmd_event_t>(e);
  });
}

int kmd_event_record(kmd_event_t event, kmd_stream_t stream) {
  return guarded([&] {
    KMD_HIP_CHECK(
        hipEventRecord(reinterpret_cast<hipEvent_t>(event), as_stream(stream)));
  });
}

int kmd_event_synchronize(kmd_event_t event) {
  return guarded([&] {
    KMD_HIP_CHECK(hipEventSynchronize(reinterpret_cast<hipEvent_t>(event)));
  });
}

int kmd_event_query(kmd_event_t event, int *done) {
  return guarded([&] {
    hipError_t e = hipEventQuery(reinterpret_cast<hipEvent_t>(event));
    if (e == hipErrorNotReady) {
      *done = 0;
      return;
    }
    KMD_HIP_CHECK(e);
    *done = 1;
  });
}

int kmd_event_elapsed_ms(kmd_event_t start, kmd_event_t stop, float *ms) {
  return guarded([&] {
    KMD_HIP_CHECK(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start),
                                      reinterpret_cast<hipEvent_t>(stop)));
  });
}

int kmd_event_destroy(kmd_event_t event) {
  return guarded([&] {
    KMD_HIP_CHECK(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)));
  });
}

int kmd_stream_wait_event(kmd_stream_t stream, kmd_event_t event) {
  return guarded([&] {
    KMD_HIP_CHECK(hipStreamWaitEvent(as_stream(stream),
                                     reinterpret_cast<hipEvent_t>(event), 0));
  });
}

int kmd_stream_synchronize(kmd_stream_t stream) {
  return guarded([&] { KMD_HIP_CHECK(hipStreamSynchronize(as_stream(stream))); });
}

int kmd_stream_create(kmd_stream_t *stream) {
  return guarded([&] {
    hipStream_t s;
    KMD_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = reinterpret_cast<kmd_stream_t>(s);
  });
}

int kmd_stream_destroy(kmd_stream_t stream) {
  return guarded([&] { KMD_HIP_CHECK(hipStreamDestroy(as_stream(stream))); });
}

int kmd_broadcast_divide(int dtype, int64_t n_rows, int64_t n_cols,
                         int64_t work_group_size, void *divident,
                         const void *divisor, kmd_stream_t stream) {
  return guarded([&] {
    check_dtype(dtype);
    require(n_rows >= 0 && n_cols >= 0, "negative array dimension");
    (void)work_group_size;
    check_ptr(divident, n_rows * n_cols, "divident");
    check_ptr(divisor, n_cols, "divisor");
    dispatch_t(dtype, [&](auto *tp) {
      using T = std::remove_pointer_t<decltype(tp)>;
      launch_broadcast_divide<T>(n_rows, n_cols, (T *)divident,
                                 (const T *)divisor, as_stream(stream));
    });
  });
}

int kmd_half_l2_norm_squared(int dtype, int64_t n_features, int64_t n_clusters,
                             int64_t work_group_size, const void *centroids,
                             void *out, kmd_stream_t stream) {
  return guarded([&] {
    check_dtype(dtype);
    require(n_features >= 0 && n_clusters >= 0, "negative array dimension");
    (void)work_group_size;
    check_ptr(centroids, n_features * n_clusters, "centroids");
    check_ptr(out, n_clusters, "out");
    dispatch_t(dtype, [&](auto *tp) {
      using T = std::remove_pointer_t<decltype(tp)>;
      launch_half_l2_norm<T>(n_features, n_clusters, (const T *)centroids,
                             (T *)out, as_stream(stream));
    });
  });
}

int kmd_reduce_centroids_data(int dtype, int itype, int64_t n_copies,
                              int64_t n_features, int64_t n_clusters,
                              int64_t work_group_size,
                              const void *cluster_sizes_private_copies,
                              const void *centroids_t_private_copies,
                              void *out_cluster_sizes, void *out_centroids_t,
                              void *out_empty_clusters_list,
                              void *out_n_empty_clusters, kmd_stream_t stream) {
  return guarded([&] {
    check_types(dtype, itype);
    require(n_copies >= 0 && n_features >= 0 && n_clusters >= 0,
            "negative array dimension");
    (void)work_group_size;
    check_ptr(cluster_sizes_private_copies, n_copies * n_clusters,
              "cluster_sizes_private_copies");
    check_ptr(centroids_t_private_copies, n_copies * n_features * n_clusters,
              "centroids_t_private_copies");
    check_ptr(out_cluster_sizes, n_clusters, "out_cluster_sizes");
    check_ptr(out_centroids_t, n_features * n_clusters, "out_centroids_t");
    check_ptr(out_empty_clusters_list, n_clusters, "out_empty_clusters_list");
    check_ptr(out_n_empty_clusters, 1, "out_n_empty_clusters");
    dispatch_ti(dtype, itype, [&](auto *tp, auto *ip) {
      using T = std::remove_pointer_t<decltype(tp)>;
      using I = std::remove_pointer_t<decltype(ip)>;
      launch_reduce_centroids<T, I>(
          n_copies, n_features, n_clusters,
          (const T *)cluster_sizes_private_copies,
          (const T *)centroids_t_private_copies, (T *)out_cluster_sizes,
          (T *)out_centroids_t, (I *)out_empty_clusters_list,
          (I *)out_n_empty_clusters, n_clusters, as_stream(stream));
    });
  });
}

int kmd_compute_threshold(int dtype, int64_t n, const void *data, int64_t topk,
                          void *threshold, kmd_stream_t stream) {
  return guarded([&] {
    check_dtype(dtype);
    require(n >= 0 && topk >= 0, "negative size");
    if (topk == 0) return;
    require(topk <= n, "topk must not exceed the number of elements");
    check_ptr(data, n, "data");
    check_ptr(threshold, 1, "threshold");
    dispatch_t(dtype, [&](auto *tp) {
      using T = std::remove_pointer_t<decltype(tp)>;
      launch_threshold<T>(n, (const T *)data, topk, (T *)threshold,
                          as_stream(stream));
    });
  });
}

int kmd_select_samples_far_from_centroid(
    int dtype, int itype, int64_t n_selected, int64_t n_samples,
    int64_t work_group_size, const void *distance_to_centroid,
    const void *threshold, void *selected_samples_idx,
    void *n_selected_gt_threshold, void *n_selected_eq_threshold,
    kmd_stream_t stream) {
  return guarded([&] {
    check_types(dtype, itype);
    require(n_selected > 0, "Argument `n_selected` must be positive");
    require(n_samples >= n_selected, "Argument `n_select` is too large");
    (void)work_group_size;
    check_ptr(distance_to_centroid, n_samples, "distance_to_centroid");
    check_ptr(selected_samples_idx, n_samples, "selected_samples_idx");
    check_ptr(threshold, 1, "threshold");
    check_ptr(n_selected_gt_threshold, 1, "n_selected_gt_threshold");
    check_ptr(n_selected_eq_threshold, 1, "n_selected_eq_threshold");
    dispatch_ti(dtype, itype, [&](auto *tp, auto *ip) {
      using T = std::remove_pointer_t<decltype(tp)>;
      using I = std::remove_pointer_t<decltype(ip)>;
      launch_select_far<T, I>(n_selected, n_samples,
                              (const T *)distance_to_centroid,
                              (const T *)threshold, (I *)selected_samples_idx,
                              (I *)n_selected_gt_threshold,
                              (I *)n_selected_eq_threshold, as_stream(stream));
    });
  });
}

int kmd_relocate_empty_clusters(
    int dtype, int itype, int64_t n_samples, int64_t n_features,
    int64_t n_clusters, int64_t n_empty_clusters, int64_t work_group_size,
    const void *X_t, const void *sample_weights, const void *assignment_id,
    const void *empty_clusters_list, const void *sq_dist_to_nearest_centroid,
    void *centroid_t, void *cluster_sizes, void *per_sample_inertia,
    kmd_stream_t stream) {
  return guarded([&] {
    check_types(dtype, itype);
    require(n_empty_clusters != 0, "n_empty_clusters must be non-zero.");
    require(n_empty_clusters > 0 && n_empty_clusters <= n_clusters,
            "n_empty_clusters must not exceed the number of clusters");
    require(n_samples > 0 && n_features > 0 && n_clusters > 0,
            "Input dimensions are inconsistent");
    (void)work_group_size;
    check_ptr(X_t, n_samples * n_features, "X_t");
    check_ptr(sample_weights, n_samples, "sample_weights");
    check_ptr(assignment_id, n_samples, "assignment_id");
    check_ptr(empty_clusters_list, n_empty_clusters, "empty_clusters_list");
    check_ptr(sq_dist_to_nearest_centroid, n_samples,
              "sq_dist_to_nearest_centroid");
    check_ptr(centroid_t, n_features * n_clusters, "centroid_t");
    check_ptr(cluster_sizes, n_clusters, "cluster_sizes");
    check_ptr(per_sample_inertia, n_samples, "per_sample_inertia");
    dispatch_ti(dtype, itype, [&](auto *tp, auto *ip) {
      using T = std::remove_pointer_t<decltype(tp)>;
      using I = std::remove_pointer_t<decltype(ip)>;
      launch_relocate<T, I>(n_samples, n_features, n_clusters, n_empty_clusters,
                            (const T *)X_t, (const T *)sample_weights,
                            (const I *)assignment_id,
                            (const I *)empty_clusters_list,
                            (const T *)sq_dist_to_nearest_centroid,
                            (T *)centroid_t, (T *)cluster_sizes,
                            (T *)per_sample_inertia, as_stream(stream));
    });
  });
}

int kmd_compute_centroid_shifts_squared(int dtype, int64_t n_features,
                                        int64_t n_clusters,
                                        const void *old_centroid_t,
                                        const void *new_centroid_t,
                                        void *out_centroid_shifts,
                                        kmd_stream_t stream) {
  return guarded([&] {
    check_dtype(dtype);
    require(n_features >= 0 && n_clusters >= 0, "negative array dimension");
    check_ptr(old_centroid_t, n_features * n_clusters, "old_centroid_t");
    check_ptr(new_centroid_t, n_features * n_clusters, "new_centroid_t");
    check_ptr(out_centroid_shifts, n_clusters, "out_centroid_shifts");
    dispatch_t(dtype, [&](auto *tp) {
      using T = std::remove_pointer_t<decltype(tp)>;
      launch_centroid_shifts<T>(n_features, n_clusters,
                                (const T *)old_centroid_t,
                                (const T *)new_centroid_t,
                                (T *)out_centroid_shifts, as_stream(stream));
    });
  });
}

int kmd_compute_centroid_to_sample_distances(
    int dtype, int64_t n_samples, int64_t n_features, int64_t n_clusters,
    int64_t work_group_size, int64_t centroids_window_height, const void *X_t,
    const void *centroid_t, void *euclidean_distances_t, kmd_stream_t stream) {
  return guarded([&] {
    check_dtype(dtype);
    require(n_samples >= 0 && n_features >= 0 && n_clusters >= 0,
            "negative array dimension");
    require(work_group_size > 0 && centroids_window_height > 0,
            "work_group_size and centroids_window_height must be positive");
    check_ptr(X_t, n_samples * n_features, "X_t");
    check_ptr(centroid_t, n_features * n_clusters, "centroid_t");
    check_ptr(euclidean_distances_t, n_clusters * n_samples,
              "euclidean_distances_t");
    if (n_features == 0 || n_clusters == 0 || n_samples == 0) {
      if (n_features == 0 && n_clusters > 0 && n_samples > 0) {
        KMD_HIP_CHECK(hipMemsetAsync(euclidean_distances_t, 0,
                                     (dtype == KMD_F32 ? 4 : 8) * n_clusters *
                                         n_samples,
                                     as_stream(stream)));
      }
      return;
    }
    dispatch_t(dtype, [&](auto *tp) {
      using T = std::remove_pointer_t<decltype(tp)>;
      launch_distances<T>(n_samples, n_features, n_clusters,
                          centroids_window_height, (const T *)X_t,
                          (const T *)centroid_t, (T *)euclidean_distances_t,
                          as_stream(stream));
    });
  });
}

int kmd_assignment(int dtype, int itype, int64_t n_samples, int64_t n_features,
                   int64_t n_clusters, int64_t centroids_window_height,
                   int64_t work_group_size, const void *X_t,
                   const void *centroids_t, const void *centroids_half_l2_norm,
                   void *assignment_id, kmd_stream_t stream) {
  return guarded([&] {
    check_types(dtype, itype);
    require(n_samples >= 0 && n_features >= 1 && n_clusters >= 1,
            "Inputs have inconsistent dimensions.");
    require(work_group_size > 0 && centroids_window_height > 0,
            "work_group_size and centroids_window_height must be positive");
    require(n_clusters <= INT32_MAX && n_features <= INT32_MAX,
            "n_clusters and n_features must fit in 32 bits");
    check_ptr(X_t, n_samples * n_features, "X_t");
    check_ptr(centroids_t, n_features * n_clusters, "centroids_t");
    check_ptr(centroids_half_l2_norm, n_clusters, "centroids_half_l2_norm");
    check_ptr(assignment_id, n_samples, "assignment_id");
    dispatch_ti(dtype, itype, [&](auto *tp, auto *ip) {
      using T = std::remove_pointer_t<decltype(tp)>;
      using I = std::remove_pointer_t<decltype(ip)>;
      launch_assignment<T, I>(n_samples, n_features, n_clusters,
                              centroids_window_height, (const T *)X_t,
                              (const T *)centroids_t,
                              (const T *)centroids_half_l2_norm,
                              (I *)assignment_id, as_stream(stream));
    });
  });
}

int kmd_compute_inertia(int dtype, int itype, int64_t n_samples,
                        int64_t n_features, int64_t n_clusters,
                        int64_t work_group_size, const void *X_t,
                        const void *sample_weight, const void *centroid_t,
                        const void *assignment_id, void *out_per_sample_inertia,
                        kmd_stream_t stream) {
  return guarded([&] {
    check_types(dtype, itype);
    require(n_samples >= 0 && n_features >= 0 && n_clusters >= 0,
            "Array dimensions are not consistent");
    (void)work_group_size;
    check_ptr(X_t, n_samples * n_features, "X_t");
    check_ptr(centroid_t, n_features * n_clusters, "centroid_t");
    check_ptr(assignment_id, n_samples, "assignment_id");
    check_ptr(out_per_sample_inertia, n_samples, "out_per_sample_inertia");
    dispatch_ti(dtype, itype, [&](auto *tp, auto *ip) {
      using T = std::remove_pointer_t<decltype(tp)>;
      using I = std::remove_pointer_t<decltype(ip)>;
      launch_inertia<T, I>(n_samples, n_features, n_clusters, (const T *)X_t,
                           (const T *)sample_weight, (const T *)centroid_t,
                           (const I *)assignment_id,
                           (T *)out_per_sample_inertia, as_stream(stream));
    });
  });
}

int kmd_reduce_vector_blocking(int dtype, int64_t n, const void *data,
                               double *result, kmd_stream_t stream) {
  return guarded([&] {
    check_dtype(dtype);
    require(n >= 0, "negative size");
    require(result != nullptr, "result must not be null");
    check_ptr(data, n, "data");
    hipStream_t s = as_stream(stream);
    TempBuffer dev(sizeof(double), s);
    dispatch_t(dtype, [&](auto *tp) {
      using T = std::remove_pointer_t<decltype(tp)>;
      if (n == 0) {
        KMD_HIP_CHECK(hipMemsetAsync(dev.as<double>(), 0, sizeof(double), s));
      } else {
        launch_sum_to_scalar<T>(n, (const T *)data, dev.as<T>(), s);
      }
      T host = T(0);
      KMD_HIP_CHECK(hipMemcpyAsync(&host, dev.as<T>(), sizeof(T),
                                   hipMemcpyDeviceToHost, s));
      KMD_HIP_CHECK(hipStreamSynchronize(s));
      *result = (double)host;
    });
  });
}

int kmd_fused_lloyd_single_step(
    int dtype, int itype, int64_t n_samples, int64_t n_features,
    int64_t n_clusters, int64_t n_copies, int64_t centroids_window_height,
    int64_t work_group_size, const void *X_t, const void *sample_weight,
    const void *centroids_t, const void *centroids_half_l2_norm,
    void *assignments_idx, void *new_centroids_t_private_copies,
    void *cluster_sizes_private_copies, kmd_stream_t stream) {
  return guarded([&] {
    check_types(dtype, itype);
    require(n_samples >= 0 && n_features >= 1 && n_clusters >= 1,
            "Unexpected array dimensions");
    require(n_copies >= 1, "number of private copies must be positive");
    require(work_group_size > 0 && centroids_window_height > 0,
            "work_group_size and centroids_window_height must be positive");
    require(n_clusters <= INT32_MAX && n_features <= INT32_MAX,
            "n_clusters and n_features must fit in 32 bits");
    check_ptr(X_t, n_samples * n_features, "X_t");
    check_ptr(sample_weight, n_samples, "sample_weight");
    check_ptr(centroids_t, n_features * n_clusters, "centroids_t");
    check_ptr(centroids_half_l2_norm, n_clusters, "centroids_half_l2_norm");
    check_ptr(assignments_idx, n_samples, "assignments_idx");
    check_ptr(new_centroids_t_private_copies, n_copies * n_features * n_clusters,
              "new_centroids_t_private_copies");
    check_ptr(cluster_sizes_private_copies, n_copies * n_clusters,
              "cluster_sizes_private_copies");
    dispatch_ti(dtype, itype, [&](auto *tp, auto *ip) {
      using T = std::remove_pointer_t<decltype(tp)>;
      using I = std::remove_pointer_t<decltype(ip)>;
      launch_fused_step_copies<T, I>(
          n_samples, n_features, n_clusters, n_copies, centroids_window_height,
          (const T *)X_t, (const T *)sample_weight, (const T *)centroids_t,
          (const T *)centroids_half_l2_norm, (I *)assignments_idx,
          (T *)new_centroids_t_private_copies, (T *)cluster_sizes_private_copies,
          as_stream(stream));
    });
  });
}

int kmd_compute_number_of_private_copies(int dtype, int device,
                                         int64_t n_samples, int64_t n_features,
                                         int64_t n_clusters, double occupancy,
                                         int64_t work_group_size,
                                         int64_t *out) {
  return guarded([&] {
    check_dtype(dtype);
    require(occupancy > 0.0 && occupancy < 1.0,
            "Expecting a fraction strictly between 0 and 1");
    require(work_group_size > 0, "work_group_size must be positive");
    require(n_samples >= 0 && n_features >= 0 && n_clusters > 0,
            "n_clusters must be positive");
    require(out != nullptr, "out must not be null");
    const int64_t cache = device_l2_bytes(device);
    const int64_t elem = dtype == KMD_F32 ? 4 : 8;
    const int64_t cluster_bytes = elem * n_clusters * (n_features + 1);
    int64_t copies =
        (int64_t)std::floor((double)cache * occupancy / (double)cluster_bytes);
    const int64_t global = ceil_div<int64_t>(n_samples, work_group_size) *
                           work_group_size;
    const int64_t subgroups = global / 8;
    if (subgroups < copies) copies = subgroups;
    if (copies < 1) copies = 1;
    *out = copies;
  });
}

int kmd_kmeans_lloyd_driver(
    int dtype, int itype, int64_t n_samples, int64_t n_features,
    int64_t n_clusters, const void *X_t, const void *sample_weight,
    void *init_centroids_t, void *assignments_id, void *res_centroids_t,
    double tol, int verbose, int64_t max_iter, int64_t centroids_window_height,
    int64_t work_group_size, double occupancy, int64_t *n_iters,
    double *total_inertia, kmd_print_fn print_fn, void *print_user_data,
    kmd_stream_t stream) {
  return guarded([&] {
    check_types(dtype, itype);
    require(occupancy > 0.0 && occupancy < 1.0,
            "Fraction `centroids_private_copies_max_cache_occupancy` is out "
            "of bounds");
    require(tol >= 0.0, "Tolerance must be non-negative");
    require(max_iter >= 0, "max_iter must be non-negative");
    require(n_samples >= 0 && n_features >= 1 && n_clusters >= 1,
            "Array dimensions are not consistent");
    require(work_group_size > 0 && centroids_window_height > 0,
            "work_group_size and centroids_window_height must be positive");
    require(n_clusters <= INT32_MAX && n_features <= INT32_MAX,
            "n_clusters and n_features must fit in 32 bits");
    require(n_iters != nullptr && total_inertia != nullptr,
            "output scalars must not be null");
    check_fit_arrays(n_samples, n_features, n_clusters, X_t, sample_weight,
                     init_centroids_t, assignments_id, res_centroids_t);
    dispatch_ti(dtype, itype, [&](auto *tp, auto *ip) {
      using T = std::remove_pointer_t<decltype(tp)>;
      using I = std::remove_pointer_t<decltype(ip)>;
      driver_lloyd<T, I>(n_samples, n_features, n_clusters, (const T *)X_t,
                         (const T *)sample_weight, (T *)init_centroids_t,
                         (I *)assignments_id, (T *)res_centroids_t, tol,
                         verbose != 0, max_iter, n_iters, total_inertia,
                         print_fn, print_user_data, as_stream(stream),
                         KMD_PREC_FP32, nullptr);
    });
  });
}

int kmd_iterative_merge_sort(int dtype, int64_t n, const void *input,
                             void *output, kmd_stream_t stream) {
  return guarded([&] {
    check_dtype(dtype);
    require(n >= 0, "negative size");
    check_ptr(input, n, "input");
    check_ptr(output, n, "output");
    dispatch_t(dtype, [&](auto *tp) {
      using T = std::remove_pointer_t<decltype(tp)>;
      launch_merge_sort<T>(n, (const T *)input, (T *)output, as_stream(stream));
    });
  });
}

int kmd_kmeans_lloyd_driver_ex(
    int dtype, int itype, int precision, int64_t n_samples,
    int64_t n_features, int64_t n_clusters, const void *X_t,
    const void *sample_weight, void *init_centroids_t, void *assignments_id,
    void *res_centroids_t, double tol, int verbose, int64_t max_iter,
    int64_t *n_iters, double *total_inertia, int64_t *n_rescored,
    double *fused_kernel_ms, int64_t *fused_launches, int *fused_path,
    kmd_print_fn print_fn, void *print_user_data, kmd_stream_t stream) {
  return guarded([&] {
    check_types(dtype, itype);
    require(precision == KMD_PREC_FP32 || precision == KMD_PREC_BF16X3 ||
                precision == KMD_PREC_BF16 || precision == KMD_PREC_FP8,
            "precision must be KMD_PREC_FP32, KMD_PREC_BF16X3, KMD_PREC_BF16 or "
            "KMD_PREC_FP8");
    require(tol >= 0.0, "Tolerance must be non-negative");
    require(max_iter >= 0, "max_iter must be non-negative");
    require(n_samples >= 0 && n_features >= 1 && n_clusters >= 1,
            "Array dimensions are not consistent");
    require(n_clusters <= INT32_MAX && n_features <= INT32_MAX,
            "n_clusters and n_features must fit in 32 bits");
    require(n_iters != nullptr && total_inertia != nullptr,
            "output scalars must not be null");
    check_fit_arrays(n_samples, n_features, n_clusters, X_t, sample_weight,
                     init_centroids_t, assignments_id, res_centroids_t);
    DriverStats st;
    dispatch_ti(dtype, itype, [&](auto *tp, auto *ip) {
      using T = std::remove_pointer_t<decltype(tp)>;
      using I = std::remove_pointer_t<decltype(ip)>;
      driver_lloyd<T, I>(n_samples, n_features, n_clusters, (const T *)X_t,
                         (const T *)sample_weight, (T *)init_centroids_t,
                         (I *)assignments_id, (T *)res_centroids_t, tol,
                         verbose != 0, max_iter, n_iters, total_inertia,
                         print_fn, print_user_data, as_stream(stream),
                         precision, &st);
    });
    if (n_rescored) *n_rescored = st.n_rescored;
    if (fused_path) *fused_path = st.fused_path;
    if (fused_kernel_ms) *fused_kernel_ms = st.fused_kernel_ms;
    if (fused_launches) *fused_launches = st.fused_launches;
  });
}

int kmd_kmeans_lloyd_batched(int64_t n_fits, kmd_fit *fits, int precision,
                             int graph_replays, int64_t *graph_launches,
                             int *assign_shape,
                             kmd_stream_t stream) {
  return guarded([&] {
    require(n_fits >= 0, "n_fits must be non-negative");
    require(n_fits == 0 || fits != nullptr, "fits must not be null");
    require(precision == KMD_PREC_FP32 || precision == KMD_PREC_BF16,
            "batched fits support precision KMD_PREC_FP32 or KMD_PREC_BF16");
    std::vector<BatchedFit> batch;
    std::vector<int64_t> idx;
    for (int64_t i = 0; i < n_fits; ++i) {
      kmd_fit &f = fits[i];
      require(f.n_samples >= 0 && f.n_features >= 1 && f.n_clusters >= 1,
              "Array dimensions are not consistent");
      require(f.tol >= 0.0, "Tolerance must be non-negative");
      require(f.max_iter >= 0, "max_iter must be non-negative");
      check_fit_arrays(f.n_samples, f.n_features, f.n_clusters, f.X_t,
                       f.sample_weight, f.init_centroids_t, f.assignments_id,
                       f.res_centroids_t);
      if (batched_fit_supported(f.n_samples, f.n_features, f.n_clusters)) {
        BatchedFit b;
        b.n = f.n_samples;
        b.d = f.n_features;
        b.k = f.n_clusters;
        b.X = (const float *)f.X_t;
        b.w = (const float *)f.sample_weight;
        b.init = (float *)f.init_centroids_t;
        b.res = (float *)f.res_centroids_t;
        b.labels = (int *)f.assignments_id;
        b.tol = f.tol;
        b.max_iter = f.max_iter;
        b.out_iters = &f.n_iters;
        b.out_inertia = &f.inertia;
        batch.push_back(b);
      } else {
        idx.push_back(i);
      }
    }
    BatchedStats st;
    const hipStream_t s = as_stream(stream);
    if (!batch.empty())
      batched_lloyd(batch, graph_replays > 0 ? graph_replays : 4, s, &st,
                    precision == KMD_PREC_BF16);
    for (int64_t i : idx) {
      kmd_fit &f = fits[i];
      driver_lloyd<float, int32_t>(
          f.n_samples, f.n_features, f.n_clusters, (const float *)f.X_t,
          (const float *)f.sample_weight, (float *)f.init_centroids_t,
          (int32_t *)f.assignments_id, (float *)f.res_centroids_t, f.tol, false,
          f.max_iter, &f.n_iters, &f.inertia, nullptr, nullptr, s, KMD_PREC_FP32,
          nullptr);
    }
    if (graph_launches) *graph_launches = st.graph_replays;
    if (assign_shape) *assign_shape = st.assign_shape;
  });
}

struct kmd_engine {
  kmd::BatchEngine *eng = nullptr;
  std::vector<int64_t> pending;
};

int kmd_engine_create(kmd_engine_t *engine, kmd_stream_t stream) {
  return guarded([&] {
    require(engine != nullptr, "engine must not be null");
    auto *e = new kmd_engine();
    try {
      e->eng = batch_engine_create(as_stream(stream));
    } catch (...) {
      delete e;
      throw;
    }
    *engine = e;
  });
}

int kmd_engine_set_precision(kmd_engine_t engine, int precision) {
  return guarded([&] {
    require(engine != nullptr, "engine must not be null");
    require(precision == KMD_PREC_FP32 || precision == KMD_PREC_BF16,
            "engine precision must be KMD_PREC_FP32 or KMD_PREC_BF16");
    batch_engine_set_mixed(engine->eng, precision == KMD_PREC_BF16);
  });
}

int kmd_engine_destroy(kmd_engine_t engine) {
  return guarded([&] {
    if (engine == nullptr) return;
    batch_engine_destroy(engine->eng);
    delete engine;
  });
}

int kmd_engine_footprint(int64_t n_samples, int64_t n_features,
                         int64_t n_clusters, int64_t *bytes) {
  return guarded([&] {
    require(bytes != nullptr, "bytes must not be null");
    require(n_samples >= 1 && n_features >= 1 && n_clusters >= 1,
            "Array dimensions are not consistent");
    *bytes = batch_fit_footprint(n_samples, n_features, n_clusters);
  });
}

int kmd_engine_add(kmd_engine_t engine, kmd_fit *fit, int64_t *id) {
  return guarded([&] {
    require(engine != nullptr && fit != nullptr && id != nullptr,
            "engine, fit and id must not be null");
    require(fit->tol >= 0.0, "Tolerance must be non-negative");
    require(fit->max_iter >= 0, "max_iter must be non-negative");
    require(batched_fit_supported(fit->n_samples, fit->n_features, fit->n_clusters),
            "engine fits need 1 <= n_features <= 64 and 1 <= n_clusters <= "
            "min(256, n_samples)");
    check_fit_arrays(fit->n_samples, fit->n_features, fit->n_clusters, fit->X_t,
                     fit->sample_weight, fit->init_centroids_t,
                     fit->assignments_id, fit->res_centroids_t);
    BatchedFit b;
    b.n = fit->n_samples;
    b.d = fit->n_features;
    b.k = fit->n_clusters;
    b.X = (const float *)fit->X_t;
    b.w = (const float *)fit->sample_weight;
    b.init = (float *)fit->init_centroids_t;
    b.res = (float *)fit->res_centroids_t;
    b.labels = (int *)fit->assignments_id;
    b.tol = fit->tol;
    b.max_iter = fit->max_iter;
    b.out_iters = &fit->n_iters;
    b.out_inertia = &fit->inertia;
    *id = batch_engine_add(engine->eng, b);
  });
}

int kmd_engine_step(kmd_engine_t engine, int iterations, int use_graph,
                    int64_t *n_active) {
  return guarded([&] {
    require(engine != nullptr, "engine must not be null");
    require(iterations >= 0, "iterations must be non-negative");
    const int64_t a = batch_engine_step(engine->eng, iterations, use_graph != 0);
    for (int64_t id : batch_engine_take_finished(engine->eng))
      engine->pending.push_back(id);
    if (n_active) *n_active = a;
  });
}

int kmd_engine_take_finished(kmd_engine_t engine, int64_t *ids,
                             int64_t capacity, int64_t *count) {
  return guarded([&] {
    require(engine != nullptr && count != nullptr, "engine and count must not be null");
    require(capacity >= 0 && (capacity == 0 || ids != nullptr),
            "ids must hold capacity entries");
    const int64_t m = std::min<int64_t>(capacity, (int64_t)engine->pending.size());
    for (int64_t i = 0; i < m; ++i) ids[i] = engine->pending[i];
    engine->pending.erase(engine->pending.begin(), engine->pending.begin() + m);
    *count = m;
  });
}

int kmd_engine_resident_bytes(kmd_engine_t engine, int64_t *bytes) {
  return guarded([&] {
    require(engine != nullptr && bytes != nullptr, "engine and bytes must not be null");
    *bytes = batch_engine_resident_bytes(engine->eng);
  });
}

int kmd_autotune(int dtype, int itype, int precision, int64_t n_samples,
                 int64_t n_features, int64_t n_clusters, const void *X_t,
                 const void *sample_weight, const void *init_centroids_t,
                 int iters, int *best_path, int *best_blocks_per_cu,
                 int *best_shape, double *best_ms, int *cand_path,
                 int *cand_blocks_per_cu, int *cand_shape, double *cand_ms,
                 int64_t capacity, int64_t *n_candidates, kmd_stream_t stream) {
  return guarded([&] {
    check_types(dtype, itype);
    require(precision == KMD_PREC_FP32 || precision == KMD_PREC_BF16X3 ||
                precision == KMD_PREC_BF16 || precision == KMD_PREC_FP8,
            "precision must be KMD_PREC_FP32, KMD_PREC_BF16X3, KMD_PREC_BF16 or "
            "KMD_PREC_FP8");
    require(n_samples >= 1 && n_features >= 1 && n_clusters >= 1,
            "Array dimensions are not consistent");
    require(iters >= 1, "iters must be positive");
    require(capacity >= 0, "capacity must be non-negative");
    check_ptr(X_t, n_samples * n_features, "X_t");
    check_ptr(sample_weight, n_samples, "sample_weight");
    check_ptr(init_centroids_t, n_features * n_clusters, "init_centroids_t");
    std::vector<TuneResult> res;
    dispatch_ti(dtype, itype, [&](auto *tp, auto *ip) {
      using T = std::remove_pointer_t<decltype(tp)>;
      using I = std::remove_pointer_t<decltype(ip)>;
      res = autotune_lloyd<T, I>(n_samples, n_features, n_clusters, (const T *)X_t,
                                 (const T *)sample_weight,
                                 (const T *)init_centroids_t, precision, iters,
                                 as_stream(stream));
    });
    const TuneResult *best = res.empty() ? nullptr : &res[0];
    for (const TuneResult &r : res)
      if (r.ms > 0.0 && (best->ms <= 0.0 || r.ms < best->ms)) best = &r;
    if (best_path) *best_path = best ? best->path : 0;
    if (best_blocks_per_cu) *best_blocks_per_cu = best ? best->blocks_per_cu : 0;
    if (best_shape) *best_shape = best ? best->shape : 0;
    if (best_ms) *best_ms = best ? best->ms : 0.0;
    const int64_t m = std::min<int64_t>(capacity, (int64_t)res.size());
    for (int64_t i = 0; i < m; ++i) {
      if (cand_path) cand_path[i] = res[i].path;
      if (cand_blocks_per_cu) cand_blocks_per_cu[i] = res[i].blocks_per_cu;
      if (cand_shape) cand_shape[i] = res[i].shape;
      if (cand_ms) cand_ms[i] = res[i].ms;
    }
    if (n_candidates) *n_candidates = (int64_t)res.size();
  });
}

int kmd_column_mean_var(int dtype, int64_t n_samples, int64_t n_features,
                        const void *X, double *mean, double *var,
                        kmd_stream_t stream) {
  return guarded([&] {
    check_dtype(dtype);
    require(n_samples >= 1 && n_features >= 1,
            "n_samples and n_features must be positive");
    check_ptr(X, n_samples * n_features, "X");
    check_ptr(mean, n_features, "mean");
    dispatch_t(dtype, [&](auto *tp) {
      using T = std::remove_pointer_t<decltype(tp)>;
      launch_column_mean_var<T>(n_samples, n_features, (const T *)X, mean, var,
                                as_stream(stream));
    });
  });
}

int kmd_transpose(int dtype, int64_t rows, int64_t cols, const void *src,
                  const double *shift, void *dst, kmd_stream_t stream) {
  return guarded([&] {
    check_dtype(dtype);
    require(rows >= 0 && cols >= 0, "negative dimensions");
    check_ptr(src, rows * cols, "src");
    check_ptr(dst, rows * cols, "dst");
    dispatch_t(dtype, [&](auto *tp) {
      using T = std::remove_pointer_t<decltype(tp)>;
      launch_transpose_shift<T>(rows, cols, (const T *)src, shift, (T *)dst,
                                as_stream(stream));
    });
  });
}

int kmd_gather_columns(int dtype, int64_t n_samples, int64_t n_features,
                       int64_t n_indices, const void *X_t,
                       const int64_t *indices, void *out_t,
                       kmd_stream_t stream) {
  return guarded([&] {
    check_dtype(dtype);
    require(n_samples >= 1 && n_features >= 1 && n_indices >= 0,
            "Inputs have inconsistent dimensions.");
    check_ptr(X_t, n_samples * n_features, "X_t");
    check_ptr(indices, n_indices, "indices");
    check_ptr(out_t, n_indices * n_features, "out_t");
    dispatch_t(dtype, [&](auto *tp) {
      using T = std::remove_pointer_t<decltype(tp)>;
      launch_gather_columns<T>(n_samples, n_features, n_indices, (const T *)X_t,
                               indices, (T *)out_t, as_stream(stream));
    });
  });
}

int kmd_kmeans_plusplus(int dtype, int64_t n_samples, int64_t n_features,
                        int64_t n_clusters, int64_t n_local_trials,
                        const void *X_t, const void *sample_weight,
                        const double *uniforms, void *centers_t,
                        int64_t *indices, kmd_stream_t stream) {
  return guarded([&] {
    check_dtype(dtype);
    require(n_features >= 1 && n_clusters >= 1,
            "n_features and n_clusters must be positive");
    require(n_samples >= n_clusters,
            "n_samples must be at least n_clusters");
    require(n_local_trials >= 1 && n_local_trials <= kpp_max_trials(),
            "n_local_trials must be between 1 and 16");
    check_ptr(X_t, n_samples * n_features, "X_t");
    check_ptr(sample_weight, n_samples, "sample_weight");
    check_ptr(uniforms, 1, "uniforms");
    check_ptr(centers_t, n_clusters * n_features, "centers_t");
    check_ptr(indices, n_clusters, "indices");
    const int64_t nu = 1 + (n_clusters - 1) * n_local_trials;
    for (int64_t i = 0; i < nu; ++i)
      require(uniforms[i] >= 0.0 && uniforms[i] < 1.0,
              "uniforms must lie in [0, 1)");
    hipStream_t s = as_stream(stream);
    TempBuffer u_dev((size_t)nu * sizeof(double), s);
    // pageable source: the copy is complete when the stream reaches it, and
    // the caller's buffer may be released when this call returns
    KMD_HIP_CHECK(hipMemcpyAsync(u_dev.as<double>(), uniforms,
                                 (size_t)nu * sizeof(double),
                                 hipMemcpyHostToDevice, s));
    KMD_HIP_CHECK(hipStreamSynchronize(s));
    dispatch_t(dtype, [&](auto *tp) {
      using T = std::remove_pointer_t<decltype(tp)>;
      launch_kmeans_plusplus<T>(n_samples, n_features, n_clusters,
                                (int)n_local_trials, (const T *)X_t,
                                (const T *)sample_weight, u_dev.as<double>(),
                                (T *)centers_t, indices, s);
    });
  });
}

int kmd_fp8_roundtrip(int64_t n, const float *in, float *out, kmd_stream_t stream) {
  return guarded([&] {
    require(n >= 0, "negative size");
    check_ptr(in, n, "in");
    check_ptr(out, n, "out");
    launch_f8_roundtrip(n, in, out, as_stream(stream));
  });
}

int kmd_set_tuning(int path, int blocks_per_cu, int shape) {
  return guarded([&] {
    require(path >= -1 && path <= 2, "path must be -1, 0, 1 or 2");
    require(shape >= 0 && shape <= 2, "shape must be 0, 1 or 2");
    set_tuning_override(path, blocks_per_cu, shape);
  });
}

}  // extern "C"
