// Launch plans and the MFMA fast path of the Lloyd step (lloyd_fast.hip).
#pragma once
#include "common.hpp"

namespace kmd {

// Timing of the fused-step kernel collected by the driver.
struct DriverStats {
  double fused_kernel_ms = 0.0;  // summed over launches
  int64_t fused_launches = 0;
  int fused_path = 0;            // 1 = MFMA fast path, 0 = generic
};

// Launch plan of a persistent fused-step kernel: grid size, dynamic LDS and
// whether centroid sums are privatised in LDS.
struct FusedPlan {
  int blocks = 1;
  size_t lds_bytes = 0;
  bool priv = false;
};

template <typename T>
FusedPlan plan_generic_fused(int64_t n, int64_t d, int64_t k);

template <typename T, typename I>
void launch_fused_generic(int64_t n, int64_t d, int64_t k, int64_t n_copies,
                          const FusedPlan &p, bool store_flush, const T *X,
                          const T *w, const T *C, const T *hn, I *labels,
                          T *copies, T *sizes_copies, hipStream_t s);

// FP32 MFMA path: true when the (d, k) shape is served by it.
template <typename T>
bool fast_fused_supported(int64_t d, int64_t k);

// Number of blocks the MFMA fused kernel launches for n samples (so callers
// that use store_flush can size one private copy per block).
int fast_fused_blocks(int64_t n, int64_t d, int64_t k);

template <typename T, typename I>
void launch_fast_fused(int64_t n, int64_t d, int64_t k, int64_t n_copies,
                       bool store_flush, const T *X, const T *w, const T *C,
                       const T *hn, I *labels, T *copies, T *sizes_copies,
                       hipStream_t s);

template <typename T, typename I>
void launch_fast_assign(int64_t n, int64_t d, int64_t k, const T *X,
                        const T *C, const T *hn, I *labels, hipStream_t s);

}  // namespace kmd
