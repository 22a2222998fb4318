// Shared host/device helpers of the gfx950 Lloyd library: error state,
// launch helpers, dtype dispatch, float atomics, stream-ordered temporaries.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "../../include/kmeans_dpcpp_amd.h"

namespace kmd {

constexpr int kWave = 64;  // CDNA wavefront width

// Error raised by argument checks; mapped to KMD_EVALUE at the C boundary.
struct ValueError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
// Error raised by HIP failures; mapped to KMD_ERUNTIME.
struct RuntimeError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void set_last_error(const std::string &msg);

#define KMD_HIP_CHECK(expr)                                                   \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      throw ::kmd::RuntimeError(std::string(#expr) + " failed: " +           \
                                hipGetErrorString(_e) + " at " __FILE__ ":" + \
                                std::to_string(__LINE__));                    \
    }                                                                         \
  } while (0)

#define KMD_LAUNCH_CHECK() KMD_HIP_CHECK(hipGetLastError())

inline void require(bool cond, const char *msg) {
  if (!cond) throw ValueError(msg);
}

template <typename T>
__host__ __device__ inline T ceil_div(T n, T m) {
  return (n + m - 1) / m;
}

// Number of compute units of the current device (cached per device).
int device_cu_count();

// Lets `kernel` launch with up to 160 KB of dynamic LDS (the MI355X CU's
// LDS); the attribute is set once per kernel (setting it costs about a
// millisecond per call).
void allow_full_lds(const void *kernel);
// L2 cache size in bytes of `device` (the reference uses SYCL's
// global_mem_cache_size, src/lloyd_single_step.hpp:23).
int64_t device_l2_bytes(int device);

// Launch tuning shared by the persistent kernels' planners: a cap on blocks
// per CU (0 = the occupancy limit) and the path choice of the driver (0 =
// automatic, 1 = LDS-resident kernels, 2 = tiled kernels).  Set by the
// autotuner (driver.hip) or kmd_set_tuning; read when a launch is planned.
int tuned_blocks_per_cu();
int tuned_path();
// Launch shape of the LDS-resident fixed-point step (0 = 16 waves x one
// 32-sample column block per wave, 1 = 12 waves x two, 2 = 8 waves x two).
int tuned_shape();
inline int apply_blocks_cap(int per_cu) {
  const int cap = tuned_blocks_per_cu();
  return (cap > 0 && per_cu > cap) ? cap : per_cu;
}

// Stream-ordered temporary device buffer (hipMallocAsync / hipFreeAsync).
// Freed in stream order when the object is destroyed, so work enqueued on the
// same stream before destruction may still use it.
class TempBuffer {
 public:
  TempBuffer(size_t bytes, hipStream_t s) : stream_(s), bytes_(bytes) {
    if (bytes_ > 0) KMD_HIP_CHECK(hipMallocAsync(&ptr_, bytes_, stream_));
  }
  ~TempBuffer() {
    if (ptr_) (void)hipFreeAsync(ptr_, stream_);
  }
  TempBuffer(const TempBuffer &) = delete;
  TempBuffer &operator=(const TempBuffer &) = delete;
  TempBuffer(TempBuffer &&o) noexcept
      : stream_(o.stream_), bytes_(o.bytes_), ptr_(o.ptr_) {
    o.ptr_ = nullptr;
    o.bytes_ = 0;
  }
  TempBuffer &operator=(TempBuffer &&o) noexcept {
    if (this != &o) {
      if (ptr_) (void)hipFreeAsync(ptr_, stream_);
      stream_ = o.stream_;
      bytes_ = o.bytes_;
      ptr_ = o.ptr_;
      o.ptr_ = nullptr;
      o.bytes_ = 0;
    }
    return *this;
  }
  template <typename T>
  T *as() const {
    return static_cast<T *>(ptr_);
  }
  size_t bytes() const { return bytes_; }

 private:
  hipStream_t stream_;
  size_t bytes_;
  void *ptr_ = nullptr;
};

// ---- device helpers -----------------------------------------------------

// Hardware float atomics (global_atomic_add_f32/f64, ds_add_f32/f64). The
// library is built with -munsafe-fp-atomics; every buffer it adds into is
// coarse-grained device memory.
template <typename T>
__device__ inline void atomic_add(T *addr, T v) {
  atomicAdd(addr, v);
}

// Running (value, index) argmin with the reference's rule: a candidate
// replaces the current best iff value < best (strict), scanning centroids in
// ascending order (src/device_functions.hpp:139-162).  For partial results
// scanned out of order, ties go to the lower index, which is the same rule.
template <typename T>
struct ArgMin {
  T val;
  int idx;
  __device__ inline void init() {
    val = std::numeric_limits<T>::infinity();
    idx = 0;
  }
  __device__ inline void take(T v, int i) {
    bool upd = (v < val) || (v == val && i < idx);
    val = upd ? v : val;
    idx = upd ? i : idx;
  }
};

template <typename T>
__device__ inline T shfl_xor_t(T v, int mask) {
  return __shfl_xor(v, mask, kWave);
}

}  // namespace kmd

// ---- internal launchers (defined in the .hip files) ------------------------
namespace kmd {

template <typename T>
void launch_broadcast_divide(int64_t n_rows, int64_t n_cols, T *x, const T *y,
                             hipStream_t s);
template <typename T>
void launch_half_l2_norm(int64_t d, int64_t k, const T *C, T *out,
                         hipStream_t s);
template <typename T, typename I>
void launch_reduce_centroids(int64_t n_copies, int64_t d, int64_t k,
                             const T *sizes_copies, const T *cent_copies,
                             T *out_sizes, T *out_cent, I *empty_list,
                             I *n_empty, int64_t list_len, hipStream_t s,
                             bool from_zero = false);
template <typename T>
void launch_centroid_shifts(int64_t d, int64_t k, const T *old_c,
                            const T *new_c, T *out, hipStream_t s);
template <typename T, typename I>
void launch_inertia(int64_t n, int64_t d, int64_t k, const T *X, const T *w,
                    const T *C, const I *labels, T *out, hipStream_t s);
// End of a Lloyd iteration in one launch: next /= sizes (per column), shifts
// against cur, FP64 sum of the shifts into out_sum[0], half squared norms of
// the new centroids into hn_next (when not null); skipped entirely while
// gate[0] != 0 (gate may be null).  partials: finalize_step_blocks(k) doubles;
// ticket: one zero-initialised counter (left at zero).
int finalize_step_blocks(int64_t k);
template <typename T, typename I>
void launch_finalize_step(int64_t d, int64_t k, const I *gate, const T *sizes,
                          const T *cur, T *next, T *shifts, T *hn_next,
                          double *partials, unsigned *ticket, double *out_sum,
                          hipStream_t s);
template <typename T>
void launch_reduce_sum(int64_t n, const T *data, double *dev_result,
                       hipStream_t s);
template <typename T>
void launch_sum_to_scalar(int64_t n, const T *data, T *dev_result,
                          hipStream_t s);
template <typename T>
void launch_max(int64_t n, const T *data, T *out, hipStream_t s);
template <typename T>
void launch_merge_sort(int64_t n, const T *in, T *out, hipStream_t s);
template <typename T>
void launch_threshold(int64_t n, const T *data, int64_t topk, T *threshold,
                      hipStream_t s);
template <typename T, typename I>
void launch_select_far(int64_t n_selected, int64_t n, const T *dist,
                       const T *threshold, I *selected, I *n_gt, I *n_eq,
                       hipStream_t s);
template <typename T, typename I>
void launch_relocate(int64_t n, int64_t d, int64_t k, int64_t n_empty,
                     const T *X, const T *w, const I *labels,
                     const I *empty_list, const T *sq_dist, T *centroids,
                     T *sizes, T *per_sample_inertia, hipStream_t s);

// Data preparation and k-means++ seeding (seeding.hip)
template <typename T>
void launch_column_mean_var(int64_t n, int64_t d, const T *X, double *mean,
                            double *var, hipStream_t s);
template <typename T>
void launch_transpose_shift(int64_t rows, int64_t cols, const T *src,
                            const double *shift, T *dst, hipStream_t s);
template <typename T>
void launch_gather_columns(int64_t n, int64_t d, int64_t m, const T *X_t,
                           const int64_t *idx, T *out_t, hipStream_t s);
int kpp_max_trials();
template <typename T>
void launch_kmeans_plusplus(int64_t n, int64_t d, int64_t k, int trials,
                            const T *X_t, const T *w, const double *u_dev,
                            T *centers_t, int64_t *indices, hipStream_t s);

// Lloyd kernels (lloyd.hip)
template <typename T, typename I>
void launch_assignment(int64_t n, int64_t d, int64_t k, int64_t h,
                       const T *X, const T *C, const T *hn, I *labels,
                       hipStream_t s);
template <typename T>
void launch_distances(int64_t n, int64_t d, int64_t k, int64_t h,
                      const T *X, const T *C, T *D, hipStream_t s);
// The reference's fused step (accumulate-into private copies).  scratch:
// fused_step_scratch_bytes<T>(d, k) bytes of device memory for the launch's
// working data (scales or centroid tiles), or NULL to allocate per call.
template <typename T>
size_t fused_step_scratch_bytes(int64_t d, int64_t k);
template <typename T, typename I>
void launch_fused_step_copies(int64_t n, int64_t d, int64_t k,
                              int64_t n_copies, int64_t h, const T *X,
                              const T *w, const T *C, const T *hn, I *labels,
                              T *copies, T *sizes_copies, hipStream_t s,
                              void *scratch = nullptr);

}  // namespace kmd
