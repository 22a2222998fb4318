// FP64 Lloyd step on the matrix core (v_mfma_f64_16x16x4f64): labels and
// the privatised centroid sums of the reference's double-precision entry
// points (python_api/_kmeans_lloyd.cpp:1043-1101 dispatch every entry for
// float64; the step itself is src/lloyd_single_step.hpp:37-209).
//
// Layout of one 16x16x4 product: rows = 16 centroids of a tile, columns =
// 16 samples, K = 4 features.  Lane l supplies A[l % 16][l / 16] (centroid
// 16 ct + l % 16, feature 4 ks + l / 16) and B[l / 16][l % 16] (feature
// 4 ks + l / 16 of sample l % 16), and receives rows l / 16 + 4 v (v = 0..3)
// of column l % 16 (the FP64 product's output layout: one row per 16-lane
// group and register, unlike the FP32 16x16x4 product's four consecutive
// rows).  The chain of a tile starts from -hn (the C operand of its
// first product), so acc = -(hn - c.x) and the best centroid is the largest
// acc: the first maximum in (tile, row) order, so ties keep the lowest
// centroid as the reference's strict-< scan does
// (src/device_functions.hpp:139-162).  The four lanes holding one sample's
// rows merge with two exchanges.
//
// Centroid sums are FP64 in LDS (ds_add_f64) per persistent block, flushed
// once per block into a partial slot (the driver, which sums the slots in
// block order) or, like the reference, added into private copy
// block % n_copies with FP64 atomics.  FP64 matrix and vector peaks are
// equal on MI355X (78.6 TFLOP/s); the matrix core keeps the vector unit free
// for the argmin and reads each centroid operand once per 16 samples.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "common.hpp"
#include "lloyd_fast.hpp"

namespace kmd {

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kNT64 = 512;         // threads per block (8 waves)
constexpr int kS64 = 2;            // 16-sample column groups per wave
constexpr size_t kLds64 = 156 * 1024;

// k-steps of 4 features: d <= 8, 16, 32, 64
inline int ks4_bucket(int64_t d) {
  return d <= 8 ? 2 : (d <= 16 ? 4 : (d <= 32 ? 8 : (d <= 64 ? 16 : 0)));
}

struct F64Args {
  int64_t n;
  int d, k, kpad;             // kpad = k rounded up to 16
  const double *X, *w, *C, *hn;
  void *labels;
  double *copies;             // [n_copies][d * k]
  double *sizes;              // [n_copies][k]
  int64_t n_copies;
  int store_flush;            // 1: write slot blockIdx.x; 0: atomics into copies
};

size_t lds64_bytes(int ks4, int64_t d, int64_t k, bool update) {
  const int64_t kpad = (k + 15) / 16 * 16;
  return sizeof(double) * ((size_t)kpad * ks4 * 4 + (size_t)kpad +
                           (update ? (size_t)(d * k + k) : 0));
}

template <int KS4, typename I, bool UPDATE>
__global__ void __launch_bounds__(kNT64) mfma64_lloyd_krn(F64Args a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  constexpr int kWaves = kNT64 / 64;
  const int d = a.d, k = a.k, kpad = a.kpad;
  const int n_ct = kpad / 16;
  const int64_t n = a.n;
  const int64_t dk = (int64_t)d * k;
  double *Al = reinterpret_cast<double *>(lds_raw);   // [n_ct][KS4][64]
  double *Hl = Al + (int64_t)n_ct * KS4 * 64;         // [kpad]: -hn, -inf pad
  double *Ps = Hl + kpad;                              // [d k + k] sums
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, j = lane & 15;

  for (int i = tid; i < n_ct * KS4 * 64; i += kNT64) {
    const int l = i & 63, ks = (i >> 6) % KS4, ct = (i >> 6) / KS4;
    const int f = 4 * ks + (l >> 4), c = 16 * ct + (l & 15);
    Al[i] = (f < d && c < k) ? a.C[(int64_t)f * k + c] : 0.0;
  }
  for (int c = tid; c < kpad; c += kNT64)
    Hl[c] = c < k ? -a.hn[c] : -__builtin_inf();
  if (UPDATE)
    for (int64_t i = tid; i < dk + k; i += kNT64) Ps[i] = 0.0;
  __syncthreads();

  I *labels = reinterpret_cast<I *>(a.labels);
  const int64_t n_super = ceil_div<int64_t>(n, 16 * kS64);
  for (int64_t sb = (int64_t)blockIdx.x * kWaves + wave; sb < n_super;
       sb += (int64_t)gridDim.x * kWaves) {
    const int64_t s0 = sb * (16 * kS64) + j;
    double xb[KS4][kS64];
#pragma unroll
    for (int ks = 0; ks < KS4; ++ks) {
      const int f = 4 * ks + g;
#pragma unroll
      for (int q = 0; q < kS64; ++q) {
        const int64_t s = s0 + 16 * q;
        xb[ks][q] = (f < d && s < n) ? a.X[(int64_t)f * n + s] : 0.0;
      }
    }
    double bestm[kS64];
    int bestc[kS64];
#pragma unroll
    for (int q = 0; q < kS64; ++q) {
      bestm[q] = -__builtin_inf();
      bestc[q] = 0;
    }
    for (int ct = 0; ct < n_ct; ++ct) {
      const double *At = Al + (int64_t)ct * KS4 * 64 + lane;
      const double *Ht = Hl + 16 * ct + g;
      f64x4 h;
      h[0] = Ht[0];
      h[1] = Ht[4];
      h[2] = Ht[8];
      h[3] = Ht[12];
      f64x4 acc[kS64];
      const double a0 = At[0];
#pragma unroll
      for (int q = 0; q < kS64; ++q)
        acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, xb[0][q], h, 0, 0, 0);
#pragma unroll
      for (int ks = 1; ks < KS4; ++ks) {
        const double av = At[64 * ks];
#pragma unroll
        for (int q = 0; q < kS64; ++q)
          acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, xb[ks][q], acc[q], 0, 0, 0);
      }
#pragma unroll
      for (int q = 0; q < kS64; ++q) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const bool upd = acc[q][v] > bestm[q];   // NaN never wins
          bestm[q] = upd ? acc[q][v] : bestm[q];
          bestc[q] = upd ? 16 * ct + 4 * v + g : bestc[q];
        }
      }
    }
    // the four lanes of a sample hold rows g, g + 4, g + 8, g + 12 of every
    // tile
#pragma unroll
    for (int q = 0; q < kS64; ++q) {
#pragma unroll
      for (int off = 16; off <= 32; off <<= 1) {
        const double om = __shfl_xor(bestm[q], off, 64);
        const int oc = __shfl_xor(bestc[q], off, 64);
        const bool take = (om > bestm[q]) || (om == bestm[q] && oc < bestc[q]);
        bestm[q] = take ? om : bestm[q];
        bestc[q] = take ? oc : bestc[q];
      }
    }
#pragma unroll
    for (int q = 0; q < kS64; ++q) {
      const int64_t s = s0 + 16 * q;
      if (s >= n) continue;
      if (g == 0) labels[s] = (I)bestc[q];
      if (UPDATE) {
        const double ws = a.w[s];
        const int lab = bestc[q];
#pragma unroll
        for (int ks = 0; ks < KS4; ++ks) {
          const int f = 4 * ks + g;
          if (f < d) atomic_add(&Ps[(int64_t)f * k + lab], xb[ks][q] * ws);
        }
        if (g == 0) atomic_add(&Ps[dk + lab], ws);
      }
    }
  }
  if (UPDATE) {
    __syncthreads();
    const int64_t copy = a.store_flush ? (int64_t)blockIdx.x : blockIdx.x % a.n_copies;
    double *cp = a.copies + copy * dk;
    double *sp = a.sizes + copy * k;
    if (a.store_flush) {
      for (int64_t i = tid; i < dk; i += kNT64) cp[i] = Ps[i];
      for (int c = tid; c < k; c += kNT64) sp[c] = Ps[dk + c];
    } else {
      for (int64_t i = tid; i < dk; i += kNT64) {
        const double v = Ps[i];
        if (v != 0.0) atomic_add(&cp[i], v);
      }
      for (int c = tid; c < k; c += kNT64) {
        const double v = Ps[dk + c];
        if (v != 0.0) atomic_add(&sp[c], v);
      }
    }
  }
}

typedef void (*Kern64)(F64Args);

template <typename I, bool UPDATE>
Kern64 pick64(int ks4) {
  switch (ks4) {
    case 2: return &mfma64_lloyd_krn<2, I, UPDATE>;
    case 4: return &mfma64_lloyd_krn<4, I, UPDATE>;
    case 8: return &mfma64_lloyd_krn<8, I, UPDATE>;
    case 16: return &mfma64_lloyd_krn<16, I, UPDATE>;
    default: return nullptr;
  }
}

// Persistent grid: CUs x resident blocks per CU, never more blocks than
// there are 16 kS64-sample super-blocks for the waves.
template <typename I, bool UPDATE>
int grid64(int64_t n, int64_t d, int64_t k) {
  const int ks4 = ks4_bucket(d);
  Kern64 fn = pick64<I, UPDATE>(ks4);
  const size_t lds = lds64_bytes(ks4, d, k, UPDATE);
  allow_full_lds(reinterpret_cast<const void *>(fn));
  int per_cu = 0;
  KMD_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void *>(fn), kNT64, lds));
  const int64_t waves_needed = ceil_div<int64_t>(n, 16 * kS64);
  const int64_t cap = ceil_div<int64_t>(waves_needed, kNT64 / 64);
  return (int)std::max<int64_t>(
      1, std::min<int64_t>((int64_t)device_cu_count() * std::max(1, per_cu), cap));
}

template <typename I, bool UPDATE>
void launch64(int64_t n, int64_t d, int64_t k, int blocks, const double *X,
              const double *w, const double *C, const double *hn, I *labels,
              double *copies, double *sizes, int64_t n_copies, bool store_flush,
              hipStream_t s) {
  const int ks4 = ks4_bucket(d);
  F64Args a{};
  a.n = n;
  a.d = (int)d;
  a.k = (int)k;
  a.kpad = (int)((k + 15) / 16 * 16);
  a.X = X;
  a.w = w;
  a.C = C;
  a.hn = hn;
  a.labels = labels;
  a.copies = copies;
  a.sizes = sizes;
  a.n_copies = std::max<int64_t>(1, n_copies);
  a.store_flush = store_flush ? 1 : 0;
  Kern64 fn = pick64<I, UPDATE>(ks4);
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(kNT64), lds64_bytes(ks4, d, k, UPDATE), s,
                     a);
  KMD_LAUNCH_CHECK();
}

}  // namespace

bool fast64_supported(int64_t d, int64_t k, bool update) {
  const int ks4 = ks4_bucket(d);
  return ks4 > 0 && k >= 1 && k <= INT32_MAX / 16 &&
         lds64_bytes(ks4, d, k, update) <= kLds64;
}

FusedPlan plan_fast64_fused(int64_t n, int64_t d, int64_t k) {
  FusedPlan p;
  p.priv = true;
  p.lds_bytes = lds64_bytes(ks4_bucket(d), d, k, true);
  p.blocks = grid64<int32_t, true>(n, d, k);
  return p;
}

template <typename I>
void launch_fast64_fused(int64_t n, int64_t d, int64_t k, int64_t n_copies,
                         const FusedPlan &p, bool store_flush, const double *X,
                         const double *w, const double *C, const double *hn,
                         I *labels, double *copies, double *sizes_copies,
                         hipStream_t s) {
  if (n <= 0) return;
  launch64<I, true>(n, d, k, p.blocks, X, w, C, hn, labels, copies, sizes_copies,
                    n_copies, store_flush, s);
}

template <typename I>
void launch_fast64_assign(int64_t n, int64_t d, int64_t k, const double *X,
                          const double *C, const double *hn, I *labels,
                          hipStream_t s) {
  if (n <= 0) return;
  const int blocks = grid64<I, false>(n, d, k);
  launch64<I, false>(n, d, k, blocks, X, nullptr, C, hn, labels, nullptr, nullptr,
                     1, false, s);
}

template void launch_fast64_fused<int32_t>(int64_t, int64_t, int64_t, int64_t,
                                           const FusedPlan &, bool, const double *,
                                           const double *, const double *,
                                           const double *, int32_t *, double *,
                                           double *, hipStream_t);
template void launch_fast64_fused<int64_t>(int64_t, int64_t, int64_t, int64_t,
                                           const FusedPlan &, bool, const double *,
                                           const double *, const double *,
                                           const double *, int64_t *, double *,
                                           double *, hipStream_t);
template void launch_fast64_assign<int32_t>(int64_t, int64_t, int64_t, const double *,
                                            const double *, const double *, int32_t *,
                                            hipStream_t);
template void launch_fast64_assign<int64_t>(int64_t, int64_t, int64_t, const double *,
                                            const double *, const double *, int64_t *,
                                            hipStream_t);

}  // namespace kmd
