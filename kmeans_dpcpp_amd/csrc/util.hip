// Small wave-level kernels around the Lloyd step: broadcast divide, half
// squared norms, private-copy reduction with the empty-cluster list, centroid
// shifts, per-sample inertia, deterministic sum / max reductions, selection of
// far samples and empty-cluster relocation.
#include "common.hpp"
#include "lloyd_fast.hpp"

namespace kmd {

namespace {

constexpr int kBlock = 256;

__device__ inline int lane_id() { return threadIdx.x & (kWave - 1); }

// Block-wide sum of a double (256 threads); result valid in every thread.
__device__ inline double block_sum(double v, double *lds /*[kBlock/kWave]*/) {
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  const int w = threadIdx.x / kWave;
  __syncthreads();
  if (lane_id() == 0) lds[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < (int)(blockDim.x / kWave); ++i) t += lds[i];
  return t;
}

// ---- broadcast divide: x[r, c] /= y[c] (src/util_kernels.hpp:9-41) ---------
template <typename T>
__global__ void __launch_bounds__(kBlock)
    broadcast_divide_krn(int64_t total, int64_t n_cols, T *x, const T *y) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * kBlock) {
    x[i] = x[i] / y[i % n_cols];
  }
}

// ---- half squared norms (src/util_kernels.hpp:47-82) -----------------------
// Column c, rows ascending, fused multiply-add accumulation, then / 2.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    half_l2_norm_krn(int64_t d, int64_t k, const T *C, T *out) {
  int64_t c = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (c >= k) return;
  T acc = T(0);
  for (int64_t f = 0; f < d; ++f) {
    T v = C[f * k + c];
    acc = fma(v, v, acc);
  }
  out[c] = acc / T(2);
}

// ---- reduce private copies (src/util_kernels.hpp:84-161) -------------------
// out[i] = sum over copies of src[copy * stride + i].  A 1024-thread block
// owns 64 consecutive elements; its 16 wavefronts each sum the copies
// congruent to their index mod 16 (ascending), then the 16 partial sums are
// added in ascending order: a fixed order, so the result is deterministic.
constexpr int kRedElems = 64;
constexpr int kRedGroups = 16;

template <typename T>
__global__ void __launch_bounds__(kRedElems * kRedGroups)
    sum_copies_krn(int64_t n_copies, int64_t m, int64_t stride,
                   const T *__restrict__ src, T *__restrict__ out) {
  __shared__ T part[kRedGroups][kRedElems];
  const int e = threadIdx.x % kRedElems, g = threadIdx.x / kRedElems;
  const int64_t i = blockIdx.x * (int64_t)kRedElems + e;
  T acc = T(0);
  if (i < m) {
    int64_t c = g;
    for (; c + 3 * kRedGroups < n_copies; c += 4 * kRedGroups) {
      const T a0 = src[c * stride + i];
      const T a1 = src[(c + kRedGroups) * stride + i];
      const T a2 = src[(c + 2 * kRedGroups) * stride + i];
      const T a3 = src[(c + 3 * kRedGroups) * stride + i];
      acc += a0;
      acc += a1;
      acc += a2;
      acc += a3;
    }
    for (; c < n_copies; c += kRedGroups) acc += src[c * stride + i];
  }
  part[g][e] = acc;
  __syncthreads();
  if (g == 0 && i < m) {
    T t = T(0);
    for (int q = 0; q < kRedGroups; ++q) t += part[q][e];
    out[i] = t;
  }
}

// Per-block partials of the MFMA fused step, [n_blocks][d*k + k]: int64
// fixed-point sums (exact, so any summation order gives the same result),
// converted once with the feature's scale; or FP32 sums in a fixed order
// when the inputs held non-finite values.
template <typename T>
__global__ void __launch_bounds__(kRedElems * kRedGroups)
    sum_partials_fixed_krn(int64_t n_blocks, int64_t d, int64_t k,
                           const void *__restrict__ partials,
                           const float *__restrict__ scale,
                           const int *__restrict__ nonfinite,
                           T *__restrict__ out_sizes, T *__restrict__ out_cent) {
  __shared__ long long ipart[kRedGroups][kRedElems];
  __shared__ float fpart[kRedGroups][kRedElems];
  const int e = threadIdx.x % kRedElems, g = threadIdx.x / kRedElems;
  const int64_t m = d * k + k;
  const int64_t i = blockIdx.x * (int64_t)kRedElems + e;
  const bool fixed = nonfinite[0] == 0;
  long long iacc = 0;
  float facc = 0.f;
  if (i < m) {
    if (fixed) {
      const long long *p = reinterpret_cast<const long long *>(partials);
      for (int64_t c = g; c < n_blocks; c += kRedGroups) iacc += p[c * m + i];
    } else {
      const float *p = reinterpret_cast<const float *>(partials);
      for (int64_t c = g; c < n_blocks; c += kRedGroups) facc += p[c * m + i];
    }
  }
  ipart[g][e] = iacc;
  fpart[g][e] = facc;
  __syncthreads();
  if (g == 0 && i < m) {
    T v;
    if (fixed) {
      long long t = 0;
      for (int q = 0; q < kRedGroups; ++q) t += ipart[q][e];
      const int64_t f = i / k;  // == d for the counts
      v = (T)((double)t / (double)scale[f]);
    } else {
      float t = 0.f;
      for (int q = 0; q < kRedGroups; ++q) t += fpart[q][e];
      v = (T)t;
    }
    if (i < d * k)
      out_cent[i] = v;
    else
      out_sizes[i - d * k] = v;
  }
}

// Appends every cluster with size 0 to empty_list at n_empty[0]++, in
// ascending cluster order (one block; prefix sums over ballots).
template <typename T, typename I>
__global__ void __launch_bounds__(1024)
    empty_list_krn(int64_t k, const T *__restrict__ sizes, I *empty_list,
                   I *n_empty, int64_t list_len, bool from_zero) {
  __shared__ int wave_counts[1024 / kWave];
  __shared__ int64_t running;
  if (threadIdx.x == 0) running = from_zero ? 0 : (int64_t)n_empty[0];
  __syncthreads();
  const int nwaves = blockDim.x / kWave;
  const int w = threadIdx.x / kWave;
  for (int64_t c0 = 0; c0 < k; c0 += blockDim.x) {
    const int64_t c = c0 + threadIdx.x;
    const bool empty = (c < k) && (sizes[c] == T(0));
    const uint64_t mask = __ballot(empty);
    const int before_in_wave = __popcll(mask & ((1ull << lane_id()) - 1ull));
    if (lane_id() == 0) wave_counts[w] = __popcll(mask);
    __syncthreads();
    int before = 0, total = 0;
    for (int i = 0; i < nwaves; ++i) {
      before += (i < w) ? wave_counts[i] : 0;
      total += wave_counts[i];
    }
    if (empty) {
      const int64_t pos = running + before + before_in_wave;
      if (pos >= 0 && pos < list_len) empty_list[pos] = (I)c;
    }
    __syncthreads();
    if (threadIdx.x == 0) running += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) n_empty[0] = (I)running;
}

// ---- centroid shifts (src/util_kernels.hpp:478-516) ------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock)
    centroid_shifts_krn(int64_t d, int64_t k, const T *a, const T *b, T *out) {
  int64_t c = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (c >= k) return;
  T acc = T(0);
  for (int64_t f = 0; f < d; ++f) {
    T diff = a[f * k + c] - b[f * k + c];
    acc = fma(diff, diff, acc);
  }
  out[c] = acc;
}

// ---- end of a Lloyd iteration: divide, shifts, their sum -------------------
// Thread c: column c of `next` divided by sizes[c] (the broadcast divide,
// src/util_kernels.hpp:9-41), then shift[c] = sum_f (cur - next)^2 in feature
// order (src/util_kernels.hpp:478-516); the k shifts are summed in FP64 (block
// sums, then the last block adds the block sums in block order).  With a
// non-null `gate`, the kernel does nothing while gate[0] != 0: the iteration
// has empty clusters, which the host relocates (on the undivided sums) before
// running it again ungated.
template <typename T, typename I>
__global__ void __launch_bounds__(kBlock)
    finalize_step_krn(int64_t d, int64_t k, const I *gate,
                      const T *__restrict__ sizes, const T *__restrict__ cur,
                      T *__restrict__ next, T *__restrict__ shifts,
                      T *__restrict__ hn_next, double *partials,
                      unsigned *ticket, double *out_sum) {
  __shared__ double lds[kBlock / kWave];
  __shared__ bool last;
  if (gate != nullptr && gate[0] != I(0)) return;
  const int64_t c = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  double v = 0.0;
  if (c < k) {
    const T sz = sizes[c];
    T acc = T(0), h = T(0);
    constexpr int kF = 8;  // features loaded ahead of their use
    for (int64_t f0 = 0; f0 < d; f0 += kF) {
      T nv[kF], cv[kF];
#pragma unroll
      for (int j = 0; j < kF; ++j) {
        if (f0 + j < d) {
          nv[j] = next[(f0 + j) * k + c];
          cv[j] = cur[(f0 + j) * k + c];
        }
      }
#pragma unroll
      for (int j = 0; j < kF; ++j) {
        if (f0 + j < d) {
          const T q = nv[j] / sz;
          next[(f0 + j) * k + c] = q;
          const T diff = cv[j] - q;
          acc = fma(diff, diff, acc);
          h = fma(q, q, h);
        }
      }
    }
    shifts[c] = acc;
    // half squared norm of the new centroid, as half_l2_norm_krn computes it
    if (hn_next != nullptr) hn_next[c] = h / T(2);
    v = (double)acc;
  }
  const double t = block_sum(v, lds);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = t;
    __threadfence();
    last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  double a = 0.0;
  const unsigned long long *pb =
      reinterpret_cast<const unsigned long long *>(partials);
  for (int i = threadIdx.x; i < (int)gridDim.x; i += kBlock)
    a += __longlong_as_double((long long)__hip_atomic_load(
        &pb[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  const double tot = block_sum(a, lds);
  if (threadIdx.x == 0) {
    out_sum[0] = tot;
    ticket[0] = 0u;
  }
}

// ---- per-sample inertia (src/compute_inertia.hpp:11-99) --------------------
template <typename T, typename I>
__global__ void __launch_bounds__(kBlock)
    inertia_krn(int64_t n, int64_t d, int64_t k, const T *X, const T *w,
                const T *C, const I *labels, T *out) {
  for (int64_t s = blockIdx.x * (int64_t)kBlock + threadIdx.x; s < n;
       s += (int64_t)gridDim.x * kBlock) {
    int64_t c = (int64_t)labels[s];
    T acc = T(0);
    if (c >= 0 && c < k) {
      for (int64_t f = 0; f < d; ++f) {
        T diff = X[f * n + s] - C[f * k + c];
        acc = fma(diff, diff, acc);
      }
    } else {
      acc = std::numeric_limits<T>::quiet_NaN();
    }
    out[s] = (w != nullptr) ? acc * w[s] : acc;
  }
}

// The same with the centroids staged in LDS as k rows of d + 1 values (row
// c = centroid c; the odd stride puts lanes with different c mod 32 on
// different banks), for k (d + 1) values within kInertiaLds bytes: each
// sample's d centroid reads come from LDS instead of a 64-address gather per
// wave load.  Same arithmetic and order per sample as inertia_krn.
constexpr size_t kInertiaLds = 64 * 1024;
template <typename T, typename I>
__global__ void __launch_bounds__(kBlock)
    inertia_lds_krn(int64_t n, int64_t d, int64_t k, const T *X, const T *w,
                    const T *C, const I *labels, T *out) {
  extern __shared__ __align__(16) unsigned char inertia_smem[];
  T *Cs = reinterpret_cast<T *>(inertia_smem);
  const int64_t ld = d + 1;
  for (int64_t e = threadIdx.x; e < d * k; e += kBlock) {
    const int64_t f = e / k, c = e - f * k;
    Cs[c * ld + f] = C[e];
  }
  __syncthreads();
  for (int64_t s = blockIdx.x * (int64_t)kBlock + threadIdx.x; s < n;
       s += (int64_t)gridDim.x * kBlock) {
    int64_t c = (int64_t)labels[s];
    T acc = T(0);
    if (c >= 0 && c < k) {
      // up to kU feature loads in flight (one round for d <= kU); the fma
      // chain still runs over f = 0, 1, ...
      constexpr int kU = 32;
      const T *row = Cs + c * ld;
      for (int64_t f0 = 0; f0 < d; f0 += kU) {
        T xv[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) xv[u] = (f0 + u < d) ? X[(f0 + u) * n + s] : T(0);
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          if (f0 + u < d) {
            T diff = xv[u] - row[f0 + u];
            acc = fma(diff, diff, acc);
          }
        }
      }
    } else {
      acc = std::numeric_limits<T>::quiet_NaN();
    }
    out[s] = (w != nullptr) ? acc * w[s] : acc;
  }
}

// ---- deterministic two-stage sum (src/compute_inertia.hpp:104-135) ---------
// Stage 1: fixed grid, each block sums a grid-stride slice in FP64.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    sum_stage1_krn(int64_t n, const T *x, double *partials) {
  __shared__ double lds[kBlock / kWave];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    acc += (double)x[i];
  double t = block_sum(acc, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

template <typename OutT>
__global__ void __launch_bounds__(kBlock)
    sum_stage2_krn(int nb, const double *partials, OutT *out) {
  __shared__ double lds[kBlock / kWave];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += kBlock) acc += partials[i];
  double t = block_sum(acc, lds);
  if (threadIdx.x == 0) out[0] = (OutT)t;
}

// ---- max reduction (src/util_kernels.hpp:174-188); NaN is ignored ----------
template <typename T>
__global__ void __launch_bounds__(kBlock)
    max_stage_krn(int64_t n, const T *x, T *out) {
  __shared__ T lds[kBlock / kWave];
  T m = -std::numeric_limits<T>::infinity();
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    T v = x[i];
    m = (v > m) ? v : m;
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    T v = __shfl_xor(m, o, kWave);
    m = (v > m) ? v : m;
  }
  if (lane_id() == 0) lds[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    T r = lds[0];
    for (int i = 1; i < kBlock / kWave; ++i) r = (lds[i] > r) ? lds[i] : r;
    out[blockIdx.x] = r;
  }
}

// ---- select samples far from centroids (src/util_kernels.hpp:208-297) -----
// Deterministic stable compaction in three launches: per-block counts,
// scatter, counter update.  Each block owns a contiguous chunk.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    select_count_krn(int64_t n, int64_t chunk, const T *dist, const T *thr,
                     int64_t *counts /*[2*nb]*/) {
  __shared__ int64_t lds[2][kBlock / kWave];
  const T t = thr[0];
  int64_t lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
  int64_t gt = 0, eq = 0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
    T v = dist[i];
    gt += (v > t);
    eq += (v == t);
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    gt += __shfl_xor(gt, o, kWave);
    eq += __shfl_xor(eq, o, kWave);
  }
  if (lane_id() == 0) {
    lds[0][threadIdx.x / kWave] = gt;
    lds[1][threadIdx.x / kWave] = eq;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t a = 0, b = 0;
    for (int i = 0; i < kBlock / kWave; ++i) {
      a += lds[0][i];
      b += lds[1][i];
    }
    counts[2 * blockIdx.x] = a;
    counts[2 * blockIdx.x + 1] = b;
  }
}

template <typename T, typename I>
__global__ void __launch_bounds__(kBlock)
    select_scatter_krn(int64_t n, int64_t chunk, int64_t n_selected,
                       const T *dist, const T *thr, const int64_t *counts,
                       int nb, I *selected, const I *n_gt, const I *n_eq) {
  __shared__ int64_t off_gt, off_eq, tot_gt;
  __shared__ int wc[2][kBlock / kWave];
  __shared__ int64_t red[3][kBlock / kWave];
  {  // this block's offsets and the total: integer sums over the blocks' counts
    int64_t a = 0, b = 0, ta = 0;
    for (int i = threadIdx.x; i < nb; i += kBlock) {
      const int64_t cg = counts[2 * i], ce = counts[2 * i + 1];
      if (i < (int)blockIdx.x) {
        a += cg;
        b += ce;
      }
      ta += cg;
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
      a += __shfl_xor(a, o, kWave);
      b += __shfl_xor(b, o, kWave);
      ta += __shfl_xor(ta, o, kWave);
    }
    if (lane_id() == 0) {
      red[0][threadIdx.x / kWave] = a;
      red[1][threadIdx.x / kWave] = b;
      red[2][threadIdx.x / kWave] = ta;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t ra = 0, rb = 0, rt = 0;
      for (int q = 0; q < kBlock / kWave; ++q) {
        ra += red[0][q];
        rb += red[1][q];
        rt += red[2][q];
      }
      off_gt = ra;
      off_eq = rb;
      tot_gt = rt;
    }
  }
  __syncthreads();
  const T t = thr[0];
  const int64_t base_gt = (int64_t)n_gt[0];
  const int64_t base_eq = (int64_t)n_eq[0];
  const int64_t eq_limit = max((int64_t)0, n_selected - tot_gt);
  int64_t run_gt = off_gt, run_eq = off_eq;
  const int w = threadIdx.x / kWave;
  int64_t lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
  for (int64_t i0 = lo; i0 < hi; i0 += kBlock) {
    int64_t i = i0 + threadIdx.x;
    bool g = false, e = false;
    if (i < hi) {
      T v = dist[i];
      g = v > t;
      e = v == t;
    }
    uint64_t mg = __ballot(g), me = __ballot(e);
    uint64_t below = (1ull << lane_id()) - 1ull;
    if (lane_id() == 0) {
      wc[0][w] = __popcll(mg);
      wc[1][w] = __popcll(me);
    }
    __syncthreads();
    int bg = 0, be = 0, tg = 0, te = 0;
    for (int q = 0; q < kBlock / kWave; ++q) {
      bg += (q < w) ? wc[0][q] : 0;
      be += (q < w) ? wc[1][q] : 0;
      tg += wc[0][q];
      te += wc[1][q];
    }
    if (g) {
      int64_t pos = base_gt + run_gt + bg + __popcll(mg & below);
      if (pos >= 0 && pos < n) selected[pos] = (I)i;
    }
    if (e) {
      int64_t r = run_eq + be + __popcll(me & below);
      if (r < eq_limit) {
        int64_t pos = n - (base_eq + r);
        if (pos >= 0 && pos < n) selected[pos] = (I)i;
      }
    }
    run_gt += tg;
    run_eq += te;
    __syncthreads();
  }
}

template <typename I>
__global__ void __launch_bounds__(kBlock)
    select_finalize_krn(int64_t n_selected, const int64_t *counts, int nb, I *n_gt,
                        I *n_eq) {
  __shared__ int64_t red[2][kBlock / kWave];
  int64_t a = 0, b = 0;
  for (int i = threadIdx.x; i < nb; i += kBlock) {
    a += counts[2 * i];
    b += counts[2 * i + 1];
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, kWave);
    b += __shfl_xor(b, o, kWave);
  }
  if (lane_id() == 0) {
    red[0][threadIdx.x / kWave] = a;
    red[1][threadIdx.x / kWave] = b;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  a = 0;
  b = 0;
  for (int q = 0; q < kBlock / kWave; ++q) {
    a += red[0][q];
    b += red[1][q];
  }
  int64_t eq_limit = n_selected - a;
  if (eq_limit < 0) eq_limit = 0;
  n_gt[0] = (I)((int64_t)n_gt[0] + a);
  n_eq[0] = (I)((int64_t)n_eq[0] + (b < eq_limit ? b : eq_limit));
}

// ---- relocation ordering + application (src/util_kernels.hpp:302-390) -----
// order[j] = j-th farthest sample by (sq_dist desc, index asc), j < m.
// Inputs: selected[0:g] = indices with dist > T (ascending index),
// selected[n-1], selected[n-2], ... = indices with dist == T (ascending).
template <typename T, typename I>
__global__ void __launch_bounds__(1024)
    relocation_order_krn(int64_t n, int64_t m, const T *dist,
                         const I *selected, const I *n_gt, const I *n_eq,
                         int64_t *order) {
  const int64_t g = (int64_t)n_gt[0];
  const int64_t e = (int64_t)n_eq[0] - 1;  // eq counter started at 1
  for (int64_t i = threadIdx.x; i < g; i += blockDim.x) {
    int64_t si = (int64_t)selected[i];
    T di = dist[si];
    int64_t rank = 0;
    for (int64_t j = 0; j < g; ++j) {
      int64_t sj = (int64_t)selected[j];
      T dj = dist[sj];
      rank += (dj > di) || (dj == di && sj < si);
    }
    if (rank < m) order[rank] = si;
  }
  for (int64_t r = threadIdx.x; r < e; r += blockDim.x) {
    int64_t pos = g + r;
    if (pos < m) order[pos] = (int64_t)selected[n - 1 - r];
  }
}

template <typename T, typename I>
__global__ void __launch_bounds__(kBlock)
    relocate_apply_krn(int64_t n, int64_t d, int64_t k, int64_t m,
                       const T *X, const T *w, const I *labels,
                       const I *empty_list, const int64_t *order, T *C,
                       T *sizes, T *inertia) {
  int64_t f = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (f >= d) return;
  for (int64_t j = 0; j < m; ++j) {
    int64_t s = order[j];
    int64_t p = (int64_t)labels[s];
    int64_t e = (int64_t)empty_list[j];
    if (s < 0 || s >= n || p < 0 || p >= k || e < 0 || e >= k) continue;
    T ws = w[s];
    T add = X[f * n + s] * ws;
    C[f * k + p] -= add;
    C[f * k + e] = add;
    if (f == 0) {
      sizes[p] -= ws;
      sizes[e] = ws;
      inertia[s] = T(0);
    }
  }
}

template <typename T, typename I>
__global__ void init_counters_krn(I *n_gt, I *n_eq) {
  n_gt[0] = I(0);
  n_eq[0] = I(1);
}

inline int grid_for(int64_t n, int per_block, int cap) {
  int64_t g = ceil_div<int64_t>(n, per_block);
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

}  // namespace

// ---- launchers ---------------------------------------------------------------

template <typename T>
void launch_broadcast_divide(int64_t n_rows, int64_t n_cols, T *x, const T *y,
                             hipStream_t s) {
  int64_t total = n_rows * n_cols;
  if (total == 0) return;
  hipLaunchKernelGGL(broadcast_divide_krn<T>, dim3(grid_for(total, kBlock, 65536)),
                     dim3(kBlock), 0, s, total, n_cols, x, y);
  KMD_LAUNCH_CHECK();
}

template <typename T>
void launch_half_l2_norm(int64_t d, int64_t k, const T *C, T *out,
                         hipStream_t s) {
  if (k == 0) return;
  hipLaunchKernelGGL(half_l2_norm_krn<T>, dim3(ceil_div<int64_t>(k, kBlock)),
                     dim3(kBlock), 0, s, d, k, C, out);
  KMD_LAUNCH_CHECK();
}

template <typename T, typename I>
void launch_reduce_centroids(int64_t n_copies, int64_t d, int64_t k,
                             const T *sizes_copies, const T *cent_copies,
                             T *out_sizes, T *out_cent, I *empty_list,
                             I *n_empty, int64_t list_len, hipStream_t s,
                             bool from_zero) {
  const int64_t dk = d * k;
  if (dk > 0) {
    hipLaunchKernelGGL(sum_copies_krn<T>, dim3(ceil_div<int64_t>(dk, kRedElems)),
                       dim3(kRedElems * kRedGroups), 0, s, n_copies, dk, dk,
                       cent_copies, out_cent);
    KMD_LAUNCH_CHECK();
  }
  if (k > 0) {
    hipLaunchKernelGGL(sum_copies_krn<T>, dim3(ceil_div<int64_t>(k, kRedElems)),
                       dim3(kRedElems * kRedGroups), 0, s, n_copies, k, k,
                       sizes_copies, out_sizes);
    KMD_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL((empty_list_krn<T, I>), dim3(1), dim3(1024), 0, s, k,
                     (const T *)out_sizes, empty_list, n_empty, list_len,
                     from_zero);
  KMD_LAUNCH_CHECK();
}

template <typename T, typename I>
void launch_reduce_fixed_partials(int64_t n_blocks, int64_t d, int64_t k,
                                  const void *partials, const float *scale,
                                  const int *nonfinite, T *out_sizes,
                                  T *out_cent, I *empty_list, I *n_empty,
                                  int64_t list_len, hipStream_t s,
                                  bool from_zero) {
  const int64_t m = d * k + k;
  hipLaunchKernelGGL(sum_partials_fixed_krn<T>, dim3(ceil_div<int64_t>(m, kRedElems)),
                     dim3(kRedElems * kRedGroups), 0, s, n_blocks, d, k,
                     partials, scale, nonfinite, out_sizes, out_cent);
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL((empty_list_krn<T, I>), dim3(1), dim3(1024), 0, s, k,
                     (const T *)out_sizes, empty_list, n_empty, list_len,
                     from_zero);
  KMD_LAUNCH_CHECK();
}

template <typename T>
void launch_centroid_shifts(int64_t d, int64_t k, const T *a, const T *b,
                            T *out, hipStream_t s) {
  if (k == 0) return;
  hipLaunchKernelGGL(centroid_shifts_krn<T>, dim3(ceil_div<int64_t>(k, kBlock)),
                     dim3(kBlock), 0, s, d, k, a, b, out);
  KMD_LAUNCH_CHECK();
}

int finalize_step_blocks(int64_t k) {
  return (int)ceil_div<int64_t>(k > 0 ? k : 1, kBlock);
}

template <typename T, typename I>
void launch_finalize_step(int64_t d, int64_t k, const I *gate, const T *sizes,
                          const T *cur, T *next, T *shifts, T *hn_next,
                          double *partials, unsigned *ticket, double *out_sum,
                          hipStream_t s) {
  hipLaunchKernelGGL((finalize_step_krn<T, I>), dim3(finalize_step_blocks(k)),
                     dim3(kBlock), 0, s, d, k, gate, sizes, cur, next, shifts,
                     hn_next, partials, ticket, out_sum);
  KMD_LAUNCH_CHECK();
}

template <typename T, typename I>
void launch_inertia(int64_t n, int64_t d, int64_t k, const T *X, const T *w,
                    const T *C, const I *labels, T *out, hipStream_t s) {
  if (n == 0) return;
  const size_t lds = sizeof(T) * (size_t)k * (size_t)(d + 1);
  if (lds <= kInertiaLds) {
    static const bool attr_set = [] {
      KMD_HIP_CHECK(hipFuncSetAttribute(
          reinterpret_cast<const void *>(&inertia_lds_krn<T, I>),
          hipFuncAttributeMaxDynamicSharedMemorySize, (int)kInertiaLds));
      return true;
    }();
    (void)attr_set;
    // persistent blocks, each staging the centroids once: as many per CU as
    // the 160 KB of LDS hold, at most 8 (32 waves)
    const int64_t per_cu = std::max<int64_t>(1, std::min<int64_t>(8, (160 * 1024) / (int64_t)lds));
    const int64_t grid = std::min<int64_t>(ceil_div<int64_t>(n, kBlock),
                                           (int64_t)device_cu_count() * per_cu);
    hipLaunchKernelGGL((inertia_lds_krn<T, I>), dim3((unsigned)grid), dim3(kBlock), lds, s,
                       n, d, k, X, w, C, labels, out);
    KMD_LAUNCH_CHECK();
    return;
  }
  hipLaunchKernelGGL((inertia_krn<T, I>), dim3(grid_for(n, kBlock, 8192)),
                     dim3(kBlock), 0, s, n, d, k, X, w, C, labels, out);
  KMD_LAUNCH_CHECK();
}

static int sum_blocks(int64_t n) { return grid_for(n, kBlock * 16, 1024); }

template <typename T>
void launch_reduce_sum(int64_t n, const T *x, double *dev_result,
                       hipStream_t s) {
  const int nb = sum_blocks(n);
  TempBuffer partials(sizeof(double) * nb, s);
  hipLaunchKernelGGL(sum_stage1_krn<T>, dim3(nb), dim3(kBlock), 0, s, n, x,
                     partials.as<double>());
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum_stage2_krn<double>, dim3(1), dim3(kBlock), 0, s, nb,
                     partials.as<const double>(), dev_result);
  KMD_LAUNCH_CHECK();
}

template <typename T>
void launch_sum_to_scalar(int64_t n, const T *x, T *dev_result,
                          hipStream_t s) {
  const int nb = sum_blocks(n);
  TempBuffer partials(sizeof(double) * nb, s);
  hipLaunchKernelGGL(sum_stage1_krn<T>, dim3(nb), dim3(kBlock), 0, s, n, x,
                     partials.as<double>());
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum_stage2_krn<T>, dim3(1), dim3(kBlock), 0, s, nb,
                     partials.as<const double>(), dev_result);
  KMD_LAUNCH_CHECK();
}

template <typename T>
void launch_max(int64_t n, const T *x, T *out, hipStream_t s) {
  const int nb = grid_for(n, kBlock * 16, 1024);
  TempBuffer partials(sizeof(T) * nb, s);
  hipLaunchKernelGGL(max_stage_krn<T>, dim3(nb), dim3(kBlock), 0, s, n, x,
                     partials.as<T>());
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(max_stage_krn<T>, dim3(1), dim3(kBlock), 0, s, (int64_t)nb,
                     partials.as<const T>(), out);
  KMD_LAUNCH_CHECK();
}

template <typename T, typename I>
void launch_select_far(int64_t n_selected, int64_t n, const T *dist,
                       const T *thr, I *selected, I *n_gt, I *n_eq,
                       hipStream_t s) {
  const int nb = grid_for(n, kBlock * 32, 1024);
  const int64_t chunk = ceil_div<int64_t>(n, nb);
  TempBuffer counts(sizeof(int64_t) * 2 * nb, s);
  hipLaunchKernelGGL(select_count_krn<T>, dim3(nb), dim3(kBlock), 0, s, n,
                     chunk, dist, thr, counts.as<int64_t>());
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL((select_scatter_krn<T, I>), dim3(nb), dim3(kBlock), 0, s,
                     n, chunk, n_selected, dist, thr,
                     counts.as<const int64_t>(), nb, selected, n_gt, n_eq);
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(select_finalize_krn<I>, dim3(1), dim3(kBlock), 0, s,
                     n_selected, counts.as<const int64_t>(), nb, n_gt, n_eq);
  KMD_LAUNCH_CHECK();
}

template <typename T, typename I>
void launch_relocate(int64_t n, int64_t d, int64_t k, int64_t n_empty,
                     const T *X, const T *w, const I *labels,
                     const I *empty_list, const T *sq_dist, T *centroids,
                     T *sizes, T *per_sample_inertia, hipStream_t s) {
  const int64_t m = n_empty < n ? n_empty : n;
  if (m <= 0 || d <= 0) return;
  TempBuffer thr(sizeof(T), s);
  TempBuffer sel(sizeof(I) * (n + 2), s);
  TempBuffer order(sizeof(int64_t) * m, s);
  I *n_gt = sel.as<I>() + n;
  I *n_eq = n_gt + 1;
  launch_threshold<T>(n, sq_dist, m, thr.as<T>(), s);
  hipLaunchKernelGGL((init_counters_krn<T, I>), dim3(1), dim3(1), 0, s, n_gt,
                     n_eq);
  KMD_LAUNCH_CHECK();
  launch_select_far<T, I>(m, n, sq_dist, thr.as<const T>(), sel.as<I>(), n_gt,
                          n_eq, s);
  hipLaunchKernelGGL((relocation_order_krn<T, I>), dim3(1), dim3(1024), 0, s, n,
                     m, sq_dist, sel.as<const I>(), n_gt, n_eq,
                     order.as<int64_t>());
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL((relocate_apply_krn<T, I>), dim3(ceil_div<int64_t>(d, kBlock)),
                     dim3(kBlock), 0, s, n, d, k, m, X, w, labels, empty_list,
                     order.as<const int64_t>(), centroids, sizes,
                     per_sample_inertia);
  KMD_LAUNCH_CHECK();
}

// ---- explicit instantiations -----------------------------------------------
#define KMD_INST_T(T)                                                          \
  template void launch_broadcast_divide<T>(int64_t, int64_t, T *, const T *,  \
                                           hipStream_t);                      \
  template void launch_half_l2_norm<T>(int64_t, int64_t, const T *, T *,      \
                                       hipStream_t);                          \
  template void launch_centroid_shifts<T>(int64_t, int64_t, const T *,        \
                                          const T *, T *, hipStream_t);       \
  template void launch_reduce_sum<T>(int64_t, const T *, double *,            \
                                     hipStream_t);                            \
  template void launch_sum_to_scalar<T>(int64_t, const T *, T *,              \
                                        hipStream_t);                         \
  template void launch_max<T>(int64_t, const T *, T *, hipStream_t);
KMD_INST_T(float)
KMD_INST_T(double)

#define KMD_INST_TI(T, I)                                                     \
  template void launch_reduce_centroids<T, I>(                                \
      int64_t, int64_t, int64_t, const T *, const T *, T *, T *, I *, I *,    \
      int64_t, hipStream_t, bool);                                                  \
  template void launch_inertia<T, I>(int64_t, int64_t, int64_t, const T *,   \
                                     const T *, const T *, const I *, T *,    \
                                     hipStream_t);                            \
  template void launch_select_far<T, I>(int64_t, int64_t, const T *,         \
                                        const T *, I *, I *, I *,             \
                                        hipStream_t);                         \
  template void launch_relocate<T, I>(int64_t, int64_t, int64_t, int64_t,    \
                                      const T *, const T *, const I *,        \
                                      const I *, const T *, T *, T *, T *,    \
                                      hipStream_t);                           \
  template void launch_reduce_fixed_partials<T, I>(                           \
      int64_t, int64_t, int64_t, const void *, const float *, const int *,   \
      T *, T *, I *, I *, int64_t, hipStream_t, bool);                              \
  template void launch_finalize_step<T, I>(int64_t, int64_t, const I *,      \
                                           const T *, const T *, T *, T *,    \
                                           T *, double *, unsigned *,         \
                                           double *, hipStream_t);
KMD_INST_TI(float, int32_t)
KMD_INST_TI(float, int64_t)
KMD_INST_TI(double, int32_t)
KMD_INST_TI(double, int64_t)

}  // namespace kmd
