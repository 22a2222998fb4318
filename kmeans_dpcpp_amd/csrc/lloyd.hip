// Generic Lloyd tile kernels (any dtype, any d and k): fused assignment +
// centroid-sum update, labels only, and sample-to-centroid distances.
//
// One lane per sample; the block stages a [kH features][kKW centroids]
// window of centroids and kKW half-norms in LDS (all lanes read the same LDS
// word, a broadcast), and each lane keeps kKW running dot products in
// registers.  Dot products accumulate over features in ascending order with
// fused multiply-adds, pseudo-inertia is hn[c] - dot[c], and the argmin keeps
// the lowest centroid index on ties -- the arithmetic of
// _acummulate_sum_of_ops / _update_closest_centroid
// (src/device_functions.hpp:106-162) without the reference's window-shape
// preconditions (src/device_functions.hpp:27-35) or its swapped distance loops
// (src/compute_euclidean_distance.hpp:56,64).
//
// The fused kernel is persistent: each block walks a grid-stride list of
// 256-sample tiles and, when k*(d+1) values fit, privatises centroid sums
// and counts in LDS (ds_add_f32/f64) and flushes them once per block into
// private copy (block % n_copies).  This replaces the reference's n*(d+1)
// global atomics (src/lloyd_single_step.hpp:169-202) with k*(d+1) per block.
#include "common.hpp"
#include "lloyd_fast.hpp"

namespace kmd {

namespace {

constexpr int kBlock = 256;
constexpr int kKW = 32;  // centroids per window
constexpr int kH = 16;   // features per window
constexpr int64_t kPrivMaxBytes = 64 * 1024;

enum { MODE_ASSIGN = 0, MODE_FUSED = 1, MODE_DIST = 2 };

template <typename T, typename I, int MODE>
__global__ void __launch_bounds__(kBlock)
    lloyd_tile_krn(int64_t n, int d, int k, const T *__restrict__ X,
                   const T *__restrict__ w, const T *__restrict__ C,
                   const T *__restrict__ hn, I *__restrict__ labels,
                   T *__restrict__ copies, T *__restrict__ sizes_copies,
                   int64_t n_copies, int priv, int store_flush,
                   T *__restrict__ D) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *cwin = reinterpret_cast<T *>(smem_raw);  // [kH][kKW]
  T *hnw = cwin + kH * kKW;                   // [kKW]
  T *psum = hnw + kKW;                        // [d][k] sums, then [k] sizes
  const int tid = threadIdx.x;
  const int64_t dk = (int64_t)d * k;

  if (MODE == MODE_FUSED && priv) {
    for (int64_t i = tid; i < dk + k; i += kBlock) psum[i] = T(0);
    __syncthreads();
  }
  const int64_t copy = (MODE == MODE_FUSED) ? (blockIdx.x % n_copies) : 0;

  const int64_t ntiles = ceil_div<int64_t>(n, kBlock);
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t s = tile * kBlock + tid;
    const bool valid = s < n;
    ArgMin<T> best;
    best.init();
    for (int c0 = 0; c0 < k; c0 += kKW) {
      T acc[kKW];
#pragma unroll
      for (int j = 0; j < kKW; ++j) acc[j] = T(0);
      if (MODE != MODE_DIST) {
        __syncthreads();
        if (tid < kKW)
          hnw[tid] = (c0 + tid < k) ? hn[c0 + tid]
                                    : std::numeric_limits<T>::infinity();
      }
      for (int f0 = 0; f0 < d; f0 += kH) {
        __syncthreads();
        for (int i = tid; i < kH * kKW; i += kBlock) {
          const int ff = i / kKW, j = i % kKW;
          const int f = f0 + ff, c = c0 + j;
          cwin[i] = (f < d && c < k) ? C[(int64_t)f * k + c] : T(0);
        }
        __syncthreads();
        const int hmax = (d - f0) < kH ? (d - f0) : kH;
        for (int ff = 0; ff < hmax; ++ff) {
          const T x = valid ? X[(int64_t)(f0 + ff) * n + s] : T(0);
          const T *row = cwin + ff * kKW;
#pragma unroll
          for (int j = 0; j < kKW; ++j) {
            if (MODE == MODE_DIST) {
              const T diff = row[j] - x;
              acc[j] = fma(diff, diff, acc[j]);
            } else {
              acc[j] = fma(row[j], x, acc[j]);
            }
          }
        }
      }
      if (MODE == MODE_DIST) {
        if (valid) {
#pragma unroll
          for (int j = 0; j < kKW; ++j)
            if (c0 + j < k) D[(int64_t)(c0 + j) * n + s] = sqrt(acc[j]);
        }
      } else {
#pragma unroll
        for (int j = 0; j < kKW; ++j) best.take(hnw[j] - acc[j], c0 + j);
      }
    }
    if (MODE != MODE_DIST && valid) labels[s] = (I)best.idx;
    if (MODE == MODE_FUSED && valid) {
      const int lab = best.idx;
      const T ws = w[s];
      if (priv) {
        for (int f = 0; f < d; ++f)
          atomic_add(&psum[(int64_t)f * k + lab], X[(int64_t)f * n + s] * ws);
        atomic_add(&psum[dk + lab], ws);
      } else {
        T *cp = copies + copy * dk;
        for (int f = 0; f < d; ++f)
          atomic_add(&cp[(int64_t)f * k + lab], X[(int64_t)f * n + s] * ws);
        atomic_add(&sizes_copies[copy * k + lab], ws);
      }
    }
  }
  if (MODE == MODE_FUSED && priv) {
    __syncthreads();
    T *cp = copies + copy * dk;
    T *sp = sizes_copies + copy * k;
    if (store_flush) {
      for (int64_t i = tid; i < dk; i += kBlock) cp[i] = psum[i];
      for (int64_t c = tid; c < k; c += kBlock) sp[c] = psum[dk + c];
    } else {
      for (int64_t i = tid; i < dk; i += kBlock) {
        T v = psum[i];
        if (v != T(0)) atomic_add(&cp[i], v);
      }
      for (int64_t c = tid; c < k; c += kBlock) {
        T v = psum[dk + c];
        if (v != T(0)) atomic_add(&sp[c], v);
      }
    }
  }
}

template <typename T>
size_t window_lds_bytes() {
  return sizeof(T) * (kH * kKW + kKW);
}

}  // namespace

// Number of blocks the persistent fused kernel launches, and whether it
// privatises in LDS, for a given problem.  Shared with the driver so that the
// private-copy buffers it allocates match the launch.
template <typename T>
FusedPlan plan_generic_fused(int64_t n, int64_t d, int64_t k) {
  // float64 shapes whose centroids and sums fit LDS: the FP64 MFMA kernel
  if (std::is_same<T, double>::value && fast64_supported(d, k, true))
    return plan_fast64_fused(n, d, k);
  FusedPlan p;
  const int64_t priv_bytes = (int64_t)sizeof(T) * (d + 1) * k;
  p.priv = priv_bytes <= kPrivMaxBytes;
  p.lds_bytes = window_lds_bytes<T>() + (p.priv ? (size_t)priv_bytes : 0);
  int per_cu = (int)((160 * 1024) / (p.lds_bytes + 1024));
  if (per_cu > 8) per_cu = 8;
  if (per_cu < 1) per_cu = 1;
  const int64_t ntiles = ceil_div<int64_t>(n, kBlock);
  int64_t blocks = (int64_t)device_cu_count() * per_cu;
  if (!p.priv) blocks = (int64_t)device_cu_count() * 8;
  if (blocks > ntiles) blocks = ntiles;
  if (blocks < 1) blocks = 1;
  p.blocks = (int)blocks;
  return p;
}

template <typename T, typename I>
void launch_fused_generic(int64_t n, int64_t d, int64_t k, int64_t n_copies,
                          const FusedPlan &p, bool store_flush, const T *X,
                          const T *w, const T *C, const T *hn, I *labels,
                          T *copies, T *sizes_copies, hipStream_t s) {
  if (n == 0) return;
  if constexpr (std::is_same<T, double>::value) {
    if (fast64_supported(d, k, true)) {
      launch_fast64_fused<I>(n, d, k, n_copies, p, store_flush, X, w, C, hn, labels,
                             copies, sizes_copies, s);
      return;
    }
  }
  hipLaunchKernelGGL((lloyd_tile_krn<T, I, MODE_FUSED>), dim3(p.blocks),
                     dim3(kBlock), p.lds_bytes, s, n, (int)d, (int)k, X, w, C,
                     hn, labels, copies, sizes_copies, n_copies, (int)p.priv,
                     (int)store_flush, (T *)nullptr);
  KMD_LAUNCH_CHECK();
}

template <typename T>
size_t fused_step_scratch_bytes(int64_t d, int64_t k) {
  if (fast_fused_supported<T>(d, k)) return 0;  // FP32 sums in LDS, no scales
  if (std::is_same<T, float>::value && tiled_supported(d, k))
    return tiled_tiles_bytes(d, k);
  return 0;  // the generic kernel works in LDS and the caller's copies only
}

template <typename T, typename I>
void launch_fused_step_copies(int64_t n, int64_t d, int64_t k,
                              int64_t n_copies, int64_t h, const T *X,
                              const T *w, const T *C, const T *hn, I *labels,
                              T *copies, T *sizes_copies, hipStream_t s,
                              void *scratch) {
  (void)h;
  if (n == 0) return;
  if (fast_fused_supported<T>(d, k)) {
    launch_fast_fused<T, I>(n, d, k, n_copies, X, w, C, hn, labels, copies,
                            sizes_copies, s);
    return;
  }
  if constexpr (std::is_same<T, float>::value) {
    // centroids or sums beyond LDS (d <= 128): MFMA assignment with centroid
    // tiles streamed through LDS, then the update from those labels
    if (tiled_supported(d, k)) {
      launch_tiled_assign<I>(n, d, k, X, C, hn, labels, s, nullptr, scratch);
      launch_label_update<I>(n, d, k, n_copies, X, w, labels, copies, sizes_copies, s);
      return;
    }
  }
  FusedPlan p = plan_generic_fused<T>(n, d, k);
  launch_fused_generic<T, I>(n, d, k, n_copies, p, false, X, w, C, hn, labels,
                             copies, sizes_copies, s);
}

template <typename T, typename I>
void launch_assignment(int64_t n, int64_t d, int64_t k, int64_t h,
                       const T *X, const T *C, const T *hn, I *labels,
                       hipStream_t s) {
  (void)h;
  if (n == 0) return;
  if (fast_fused_supported<T>(d, k)) {
    launch_fast_assign<T, I>(n, d, k, X, C, hn, labels, s);
    return;
  }
  if constexpr (std::is_same<T, float>::value) {
    if (tiled_supported(d, k)) {
      launch_tiled_assign<I>(n, d, k, X, C, hn, labels, s);
      return;
    }
  } else {
    if (fast64_supported(d, k, false)) {
      launch_fast64_assign<I>(n, d, k, X, C, hn, labels, s);
      return;
    }
  }
  const int64_t blocks = ceil_div<int64_t>(n, kBlock);
  hipLaunchKernelGGL((lloyd_tile_krn<T, I, MODE_ASSIGN>), dim3(blocks),
                     dim3(kBlock), window_lds_bytes<T>(), s, n, (int)d, (int)k,
                     X, (const T *)nullptr, C, hn, labels, (T *)nullptr,
                     (T *)nullptr, (int64_t)1, 0, 0, (T *)nullptr);
  KMD_LAUNCH_CHECK();
}

template <typename T>
void launch_distances(int64_t n, int64_t d, int64_t k, int64_t h, const T *X,
                      const T *C, T *D, hipStream_t s) {
  (void)h;
  if (n == 0) return;
  const int64_t blocks = ceil_div<int64_t>(n, kBlock);
  hipLaunchKernelGGL((lloyd_tile_krn<T, int32_t, MODE_DIST>), dim3(blocks),
                     dim3(kBlock), window_lds_bytes<T>(), s, n, (int)d, (int)k,
                     X, (const T *)nullptr, C, (const T *)nullptr,
                     (int32_t *)nullptr, (T *)nullptr, (T *)nullptr,
                     (int64_t)1, 0, 0, D);
  KMD_LAUNCH_CHECK();
}

#define KMD_INST_LLOYD(T, I)                                                   \
  template void launch_fused_generic<T, I>(                                    \
      int64_t, int64_t, int64_t, int64_t, const FusedPlan &, bool, const T *,  \
      const T *, const T *, const T *, I *, T *, T *, hipStream_t);            \
  template void launch_fused_step_copies<T, I>(                                \
      int64_t, int64_t, int64_t, int64_t, int64_t, const T *, const T *,       \
      const T *, const T *, I *, T *, T *, hipStream_t, void *);               \
  template void launch_assignment<T, I>(int64_t, int64_t, int64_t, int64_t,   \
                                        const T *, const T *, const T *, I *,  \
                                        hipStream_t);
KMD_INST_LLOYD(float, int32_t)
KMD_INST_LLOYD(float, int64_t)
KMD_INST_LLOYD(double, int32_t)
KMD_INST_LLOYD(double, int64_t)
template size_t fused_step_scratch_bytes<float>(int64_t, int64_t);
template size_t fused_step_scratch_bytes<double>(int64_t, int64_t);
template FusedPlan plan_generic_fused<float>(int64_t, int64_t, int64_t);
template FusedPlan plan_generic_fused<double>(int64_t, int64_t, int64_t);
template void launch_distances<float>(int64_t, int64_t, int64_t, int64_t,
                                      const float *, const float *, float *,
                                      hipStream_t);
template void launch_distances<double>(int64_t, int64_t, int64_t, int64_t,
                                       const double *, const double *,
                                       double *, hipStream_t);

}  // namespace kmd
