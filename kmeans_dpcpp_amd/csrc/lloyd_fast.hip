// FP32 MFMA fast path of the fused Lloyd step and of the assignment
// (float data, d <= 64, centroids + sums fitting in LDS).
//
// The cross term dot[c, s] = sum_f C[f, c] X[f, s] is a
// (k x d) . (d x n) product computed with v_mfma_f32_32x32x2f32: the 32 rows
// of an MFMA tile are centroids, its 32 columns are samples, its K = 2 steps
// walk the features.  In that layout each lane's 16 accumulators belong to a
// single sample (column lane % 32) and 16 centroids, so the argmin is 16
// register compares plus one exchange with lane ^ 32.
//
// Per wave, a super-block of 32 * S samples is processed at once: lane l
// holds X[f][s0 + S * (l % 32) + q] for its feature parity f = 2 ks + l / 32
// and q < S, read once from HBM with one S-wide load per feature (the
// column a sample occupies is a free choice of the product, so neighbouring
// samples sit in one lane).  The whole centroid matrix (zero-padded to 32
// columns and an even number of rows) and the half-norms (+inf padded) stay
// in LDS for the life of the block, and every A operand read from LDS feeds
// S MFMAs.  The update adds w x into per-block centroid sums and counts in
// LDS (ds_add_f32), flushed once per block either as a plain store into the
// block's own private copy (driver; deterministic) or with global atomics
// into copy (block % n_copies) (the fused_lloyd_single_step entry point).
//
// Arithmetic: FP32 products accumulated in FP32 by the matrix core; the
// pseudo-inertia hn[c] - dot and the argmin rule (strict <, lowest index on
// ties) are those of src/device_functions.hpp:139-162.  Relative to the
// sequential FMA chain of the reference the dot products differ by a few
// ulps, so labels can differ only on sub-ulp margins.
#include "lloyd_fast.hpp"

namespace kmd {

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / kWave;
// 32-sample MFMA column blocks per wave (S).  S = 2 keeps the d = 32 kernel
// at 3 waves per SIMD; S = 4 halves the LDS reads of A operands.
constexpr int kSDefault = 2;

inline int mfma_s() {
  static const int s = [] {
    const char *e = std::getenv("KMD_MFMA_S");
    return (e && std::atoi(e) == 4) ? 4 : kSDefault;
  }();
  return s;
}
constexpr int kMaxD = 64;
constexpr size_t kLdsPerCu = 160 * 1024;

inline int pad32(int64_t k) { return (int)ceil_div<int64_t>(k, 32) * 32; }

// xr[ks][q] = X[2 ks + half][s_first + q] (0 outside the matrix).  FULL:
// d == 2 KS, so every feature row exists.  VEC: rows are 16-byte aligned and
// s_first is a multiple of S, so one S-wide load reads the S samples.
template <int KS, int S, bool FULL>
__device__ inline void load_x(const float *__restrict__ X, int64_t n, int d,
                              int64_t s_first, int half, bool vec,
                              float (&xr)[KS][S]) {
  const bool whole = vec && (s_first + S <= n);
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const int f = 2 * ks + half;
    if (FULL || f < d) {
      const float *row = X + (int64_t)f * n + s_first;
      if (whole) {
        if constexpr (S == 4) {
          const float4 v = *reinterpret_cast<const float4 *>(row);
          xr[ks][0] = v.x;
          xr[ks][1] = v.y;
          xr[ks][2] = v.z;
          xr[ks][3] = v.w;
        } else if constexpr (S == 2) {
          const float2 v = *reinterpret_cast<const float2 *>(row);
          xr[ks][0] = v.x;
          xr[ks][1] = v.y;
        } else {
#pragma unroll
          for (int q = 0; q < S; ++q) xr[ks][q] = row[q];
        }
      } else {
#pragma unroll
        for (int q = 0; q < S; ++q)
          xr[ks][q] = (s_first + q < n) ? row[q] : 0.f;
      }
    } else {
#pragma unroll
      for (int q = 0; q < S; ++q) xr[ks][q] = 0.f;
    }
  }
}

template <int KS, int S, bool FULL, typename I, bool UPDATE, bool STORE_FLUSH>
__global__ void __launch_bounds__(kThreads)
    mfma_lloyd_krn(int64_t n, int d, int k, int kpad, int vec,
                   const float *__restrict__ X, const float *__restrict__ w,
                   const float *__restrict__ C, const float *__restrict__ hn,
                   I *__restrict__ labels, float *__restrict__ copies,
                   float *__restrict__ sizes_copies, int64_t n_copies) {
  extern __shared__ __align__(16) float lds[];
  float *Cl = lds;                    // [2*KS][kpad] centroid rows
  float *Hl = Cl + 2 * KS * kpad;     // [kpad] half-norms
  float *Sl = Hl + kpad;              // [d][k] sums then [k] counts
  const int tid = threadIdx.x;
  const int64_t dk = (int64_t)d * k;

  for (int i = tid; i < 2 * KS * kpad; i += kThreads) {
    const int f = i / kpad, c = i - f * kpad;
    Cl[i] = (f < d && c < k) ? C[(int64_t)f * k + c] : 0.f;
  }
  for (int c = tid; c < kpad; c += kThreads)
    Hl[c] = (c < k) ? hn[c] : __builtin_huge_valf();
  if (UPDATE)
    for (int64_t i = tid; i < dk + k; i += kThreads) Sl[i] = 0.f;
  __syncthreads();

  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;
  const int col = lane & 31;
  const int half = lane >> 5;
  const int n_ct = kpad / 32;
  constexpr int kSuper = 32 * S;
  const int64_t n_super = ceil_div<int64_t>(n, kSuper);

  for (int64_t sb = (int64_t)blockIdx.x * kWavesPerBlock + wave; sb < n_super;
       sb += (int64_t)gridDim.x * kWavesPerBlock) {
    const int64_t s_first = sb * kSuper + (int64_t)S * col;
    float xr[KS][S];
    load_x<KS, S, FULL>(X, n, d, s_first, half, vec != 0, xr);

    float bestv[S];
    int besti[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
      bestv[q] = __builtin_huge_valf();
      besti[q] = 0;
    }
    for (int ct = 0; ct < n_ct; ++ct) {
      f32x16 acc[S];
#pragma unroll
      for (int q = 0; q < S; ++q)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[q][r] = 0.f;
      const float *Ccol = Cl + ct * 32 + col;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const float a = Ccol[(2 * ks + half) * kpad];
#pragma unroll
        for (int q = 0; q < S; ++q)
          acc[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, xr[ks][q], acc[q], 0,
                                                         0, 0);
      }
      // rows held by this lane: m = 8 (r / 4) + 4 half + (r % 4), ascending in r
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int m0 = ct * 32 + 8 * g + 4 * half;
        const float4 h4 = *reinterpret_cast<const float4 *>(Hl + m0);
        const float hv[4] = {h4.x, h4.y, h4.z, h4.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
#pragma unroll
          for (int q = 0; q < S; ++q) {
            const float v = hv[e] - acc[q][4 * g + e];
            const bool upd = v < bestv[q];
            bestv[q] = upd ? v : bestv[q];
            besti[q] = upd ? (m0 + e) : besti[q];
          }
        }
      }
    }
    // merge the two halves of the column (lanes l and l ^ 32)
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const float ov = __shfl_xor(bestv[q], 32, kWave);
      const int oi = __shfl_xor(besti[q], 32, kWave);
      const bool take = (ov < bestv[q]) || (ov == bestv[q] && oi < besti[q]);
      bestv[q] = take ? ov : bestv[q];
      besti[q] = take ? oi : besti[q];
    }
    if (half == 0) {
#pragma unroll
      for (int q = 0; q < S; ++q)
        if (s_first + q < n) labels[s_first + q] = (I)besti[q];
    }
    if (UPDATE) {
#pragma unroll
      for (int q = 0; q < S; ++q) {
        if (s_first + q < n) {
          const float ws = w[s_first + q];
          const int lab = besti[q];
#pragma unroll
          for (int ks = 0; ks < KS; ++ks) {
            const int f = 2 * ks + half;
            if (FULL || f < d)
              atomic_add(&Sl[(int64_t)f * k + lab], xr[ks][q] * ws);
          }
          if (half == 0) atomic_add(&Sl[dk + lab], ws);
        }
      }
    }
  }
  if (UPDATE) {
    __syncthreads();
    const int64_t copy = STORE_FLUSH ? blockIdx.x : (blockIdx.x % n_copies);
    float *cp = copies + copy * dk;
    float *sp = sizes_copies + copy * k;
    if (STORE_FLUSH) {
      for (int64_t i = tid; i < dk; i += kThreads) cp[i] = Sl[i];
      for (int64_t c = tid; c < k; c += kThreads) sp[c] = Sl[dk + c];
    } else {
      for (int64_t i = tid; i < dk; i += kThreads) {
        const float v = Sl[i];
        if (v != 0.f) atomic_add(&cp[i], v);
      }
      for (int64_t c = tid; c < k; c += kThreads) {
        const float v = Sl[dk + c];
        if (v != 0.f) atomic_add(&sp[c], v);
      }
    }
  }
}

inline int ks_bucket(int64_t d) {
  const int ks = (int)ceil_div<int64_t>(d, 2);
  if (ks <= 2) return 2;
  if (ks <= 4) return 4;
  if (ks <= 8) return 8;
  if (ks <= 12) return 12;
  if (ks <= 16) return 16;
  if (ks <= 24) return 24;
  return 32;
}

inline size_t lds_bytes(int64_t d, int64_t k, bool update) {
  const int kpad = pad32(k);
  const int KS = ks_bucket(d);
  size_t b = sizeof(float) * ((size_t)2 * KS * kpad + kpad);
  if (update) b += sizeof(float) * (size_t)(d + 1) * k;
  return b;
}

inline int blocks_per_cu(size_t lds) {
  int b = (int)(kLdsPerCu / (lds + 512));
  if (b > 4) b = 4;
  return b;
}

int plan_blocks(int64_t n, size_t lds, int S) {
  const int64_t n_super = ceil_div<int64_t>(n, 32 * S);
  int64_t blocks = (int64_t)device_cu_count() * blocks_per_cu(lds);
  const int64_t need = ceil_div<int64_t>(n_super, kWavesPerBlock);
  if (blocks > need) blocks = need;
  if (blocks < 1) blocks = 1;
  return (int)blocks;
}

template <int KS, int S, bool FULL, typename I, bool UPDATE, bool STORE>
void launch_ks(int64_t n, int64_t d, int64_t k, int blocks, size_t lds,
               bool vec, const float *X, const float *w, const float *C,
               const float *hn, I *labels, float *copies, float *sizes_copies,
               int64_t n_copies, hipStream_t s) {
  auto kern = mfma_lloyd_krn<KS, S, FULL, I, UPDATE, STORE>;
  static bool attr_set = false;  // per instantiation
  if (!attr_set) {
    KMD_HIP_CHECK(hipFuncSetAttribute((const void *)kern,
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)kLdsPerCu));
    attr_set = true;
  }
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(kThreads), lds, s, n, (int)d,
                     (int)k, pad32(k), (int)vec, X, w, C, hn, labels, copies,
                     sizes_copies, n_copies);
  KMD_LAUNCH_CHECK();
}

template <typename I, bool UPDATE, bool STORE>
void dispatch_ks(int64_t n, int64_t d, int64_t k, int blocks, size_t lds,
                 bool vec, const float *X, const float *w, const float *C,
                 const float *hn, I *labels, float *copies,
                 float *sizes_copies, int64_t n_copies, hipStream_t s) {
#define KMD_KS_LAUNCH(KS, SS)                                             \
  if (d == 2 * KS)                                                        \
    launch_ks<KS, SS, true, I, UPDATE, STORE>(n, d, k, blocks, lds, vec,  \
                                              X, w, C, hn, labels,        \
                                              copies, sizes_copies,       \
                                              n_copies, s);               \
  else                                                                    \
    launch_ks<KS, SS, false, I, UPDATE, STORE>(n, d, k, blocks, lds, vec, \
                                               X, w, C, hn, labels,       \
                                               copies, sizes_copies,      \
                                               n_copies, s);
#define KMD_KS_CASE(KS)                                                   \
  case KS:                                                                \
    if (mfma_s() == 4) {                                                  \
      KMD_KS_LAUNCH(KS, 4)                                                \
    } else {                                                              \
      KMD_KS_LAUNCH(KS, 2)                                                \
    }                                                                     \
    return;
  switch (ks_bucket(d)) {
    KMD_KS_CASE(2)
    KMD_KS_CASE(4)
    KMD_KS_CASE(8)
    KMD_KS_CASE(12)
    KMD_KS_CASE(16)
    KMD_KS_CASE(24)
    KMD_KS_CASE(32)
    default:
      throw RuntimeError("unsupported feature count for the MFMA path");
  }
#undef KMD_KS_CASE
#undef KMD_KS_LAUNCH
}

inline bool vec_ok(const float *X, int64_t n) {
  const int64_t S = mfma_s();
  return (n % S == 0) && ((reinterpret_cast<uintptr_t>(X) & (4 * S - 1)) == 0);
}

}  // namespace

template <typename T>
bool fast_fused_supported(int64_t d, int64_t k) {
  if (!std::is_same<T, float>::value) return false;
  if (d < 1 || d > kMaxD || k < 1) return false;
  return blocks_per_cu(lds_bytes(d, k, true)) >= 1;
}

int fast_fused_blocks(int64_t n, int64_t d, int64_t k) {
  return plan_blocks(n, lds_bytes(d, k, true), mfma_s());
}

template <typename T, typename I>
void launch_fast_fused(int64_t n, int64_t d, int64_t k, int64_t n_copies,
                       bool store_flush, const T *X, const T *w, const T *C,
                       const T *hn, I *labels, T *copies, T *sizes_copies,
                       hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    if (n <= 0) return;
    const size_t lds = lds_bytes(d, k, true);
    const int blocks = plan_blocks(n, lds, mfma_s());
    if (store_flush && n_copies < blocks)
      throw RuntimeError("store-flush needs one private copy per block");
    if (store_flush)
      dispatch_ks<I, true, true>(n, d, k, blocks, lds, vec_ok(X, n), X, w, C,
                                 hn, labels, copies, sizes_copies, n_copies, s);
    else
      dispatch_ks<I, true, false>(n, d, k, blocks, lds, vec_ok(X, n), X, w, C,
                                  hn, labels, copies, sizes_copies, n_copies,
                                  s);
  } else {
    throw RuntimeError("MFMA fused step is FP32 only");
  }
}

template <typename T, typename I>
void launch_fast_assign(int64_t n, int64_t d, int64_t k, const T *X,
                        const T *C, const T *hn, I *labels, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    if (n <= 0) return;
    const size_t lds = lds_bytes(d, k, false);
    const int blocks = plan_blocks(n, lds, mfma_s());
    dispatch_ks<I, false, false>(n, d, k, blocks, lds, vec_ok(X, n), X,
                                 nullptr, C, hn, labels, nullptr, nullptr, 1,
                                 s);
  } else {
    throw RuntimeError("MFMA assignment is FP32 only");
  }
}

template bool fast_fused_supported<float>(int64_t, int64_t);
template bool fast_fused_supported<double>(int64_t, int64_t);
#define KMD_INST_FAST(T, I)                                                   \
  template void launch_fast_fused<T, I>(int64_t, int64_t, int64_t, int64_t,   \
                                        bool, const T *, const T *,           \
                                        const T *, const T *, I *, T *, T *,  \
                                        hipStream_t);                         \
  template void launch_fast_assign<T, I>(int64_t, int64_t, int64_t,           \
                                         const T *, const T *, const T *,     \
                                         I *, hipStream_t);
KMD_INST_FAST(float, int32_t)
KMD_INST_FAST(float, int64_t)
KMD_INST_FAST(double, int32_t)
KMD_INST_FAST(double, int64_t)

}  // namespace kmd
