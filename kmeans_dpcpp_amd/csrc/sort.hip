// Stable ascending sort of floating-point values (NaN last) and the top-k
// threshold (the element the sort puts at n - topk, found by radix select
// over the same order keys, without sorting).
//
// Replaces iterative_merge_sort (src/iterative_merge_sort.hpp:204-368) and
// compute_threshold_kernel (src/util_kernels.hpp:163-206).  The reference sorts
// work-group tiles with sort_over_group and then runs log2(n/lws) global merge
// passes with a per-work-item binary search.  Here a 256-thread block sorts a
// 2048-element tile in LDS (8 elements per thread in registers, then 8 merge
// rounds in LDS), and every global merge pass is block-cooperative: one
// merge-path search per block locates its slice of both runs, the slice is
// staged in LDS and each thread merges 8 outputs.  All merges take the left
// run first on ties, so the sort is stable and its output is the unique
// stable ascending order (bit-exact against any stable ascending sort).
#include "common.hpp"

namespace kmd {

namespace {

constexpr int kSortBlock = 256;
constexpr int kItems = 8;
constexpr int kTile = kSortBlock * kItems;  // 2048

// Total order used by the sort: -0 == +0, every NaN compares equal to every
// other NaN and greater than all numbers.  Equal elements keep input order.
template <typename T>
struct Key;
template <>
struct Key<float> {
  using U = uint32_t;
  __host__ __device__ static inline U of(float v) {
    U u = __builtin_bit_cast(U, v);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0xffffffffu;  // NaN
    if (u == 0x80000000u) u = 0u;                             // -0 -> +0
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  }
};
template <>
struct Key<double> {
  using U = uint64_t;
  __host__ __device__ static inline U of(double v) {
    U u = __builtin_bit_cast(U, v);
    if ((u & 0x7fffffffffffffffull) > 0x7ff0000000000000ull)
      return 0xffffffffffffffffull;
    if (u == 0x8000000000000000ull) u = 0ull;
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
  }
};

template <typename T>
__device__ inline bool less(T a, T b) {
  return Key<T>::of(a) < Key<T>::of(b);
}

// Number of elements taken from A among the first `diag` outputs of the
// stable merge of A[0:na] and B[0:nb] (A first on ties).
template <typename T>
__device__ inline int64_t merge_path(const T *A, int64_t na, const T *B,
                                     int64_t nb, int64_t diag) {
  int64_t lo = diag > nb ? diag - nb : 0;
  int64_t hi = diag < na ? diag : na;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (!less(B[diag - 1 - mid], A[mid]))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Merge kItems outputs starting at A[i], B[j] into out[0:cnt].
template <typename T>
__device__ inline void merge_seq(const T *A, int64_t na, const T *B,
                                 int64_t nb, int64_t i, int64_t j, T *out,
                                 int cnt) {
  for (int q = 0; q < cnt; ++q) {
    bool takeA = (i < na) && (j >= nb || !less(B[j], A[i]));
    out[q] = takeA ? A[i] : B[j];
    i += takeA;
    j += !takeA;
  }
}

// Sort one tile of up to kTile elements; writes the sorted tile to out.
template <typename T>
__global__ void __launch_bounds__(kSortBlock)
    block_sort_krn(int64_t n, const T *in, T *out) {
  __shared__ T buf[2][kTile];
  const int64_t base = blockIdx.x * (int64_t)kTile;
  const int len = (int)((n - base) < kTile ? (n - base) : kTile);
  const T pad = std::numeric_limits<T>::quiet_NaN();
  for (int i = threadIdx.x; i < kTile; i += kSortBlock)
    buf[0][i] = (i < len) ? in[base + i] : pad;
  __syncthreads();
  // per-thread stable sort of kItems consecutive elements (odd-even
  // transposition; swaps only on strict inequality)
  T r[kItems];
  for (int q = 0; q < kItems; ++q) r[q] = buf[0][threadIdx.x * kItems + q];
#pragma unroll
  for (int pass = 0; pass < kItems; ++pass) {
#pragma unroll
    for (int q = pass & 1; q + 1 < kItems; q += 2) {
      if (less(r[q + 1], r[q])) {
        T t = r[q];
        r[q] = r[q + 1];
        r[q + 1] = t;
      }
    }
  }
  for (int q = 0; q < kItems; ++q) buf[0][threadIdx.x * kItems + q] = r[q];
  __syncthreads();
  // merge rounds in LDS: runs of L -> 2L.  Pads are NaN and sit at the end
  // of the tile, so stability keeps them after every real NaN.
  int src = 0;
  for (int L = kItems; L < kTile; L <<= 1) {
    const int o = threadIdx.x * kItems;     // first output of this thread
    const int pair = o / (2 * L);
    const int a0 = pair * 2 * L;
    const T *A = &buf[src][a0];
    const T *B = &buf[src][a0 + L];
    const int diag = o - a0;
    int i = (int)merge_path(A, L, B, L, diag);
    int j = diag - i;
    T tmp[kItems];
    merge_seq(A, L, B, L, i, j, tmp, kItems);
    for (int q = 0; q < kItems; ++q) buf[src ^ 1][o + q] = tmp[q];
    src ^= 1;
    __syncthreads();
  }
  for (int i = threadIdx.x; i < len; i += kSortBlock) out[base + i] = buf[src][i];
}

// One global merge pass: runs of length L in `in` merged pairwise into out.
template <typename T>
__global__ void __launch_bounds__(kSortBlock)
    merge_pass_krn(int64_t n, int64_t L, const T *in, T *out) {
  __shared__ T stage[kTile];
  __shared__ int64_t split[2];
  const int64_t o0 = blockIdx.x * (int64_t)kTile;  // first global output
  const int64_t pair_start = (o0 / (2 * L)) * (2 * L);   // < n
  const int64_t na = (n - pair_start) < L ? (n - pair_start) : L;
  const int64_t b_start = pair_start + L;
  const int64_t nb_raw = n - b_start;
  const int64_t nb = nb_raw < 0 ? 0 : (nb_raw > L ? L : nb_raw);
  const int64_t total = na + nb;
  const int64_t d0 = o0 - pair_start;
  const int64_t d1 = (d0 + kTile < total) ? d0 + kTile : total;
  const T *A = in + pair_start;
  const T *B = in + b_start;  // only read when nb > 0
  if (threadIdx.x < 2) {
    int64_t diag = threadIdx.x == 0 ? d0 : d1;
    split[threadIdx.x] = merge_path(A, na, B, nb, diag);
  }
  __syncthreads();
  const int64_t ia0 = split[0], ia1 = split[1];
  const int64_t jb0 = d0 - ia0, jb1 = d1 - ia1;
  const int lenA = (int)(ia1 - ia0), lenB = (int)(jb1 - jb0);
  for (int t = threadIdx.x; t < lenA; t += kSortBlock) stage[t] = A[ia0 + t];
  for (int t = threadIdx.x; t < lenB; t += kSortBlock)
    stage[lenA + t] = B[jb0 + t];
  __syncthreads();
  const int cnt_total = lenA + lenB;
  const int o = threadIdx.x * kItems;
  T tmp[kItems];
  int cnt = 0;
  if (o < cnt_total) {
    const T *sA = stage;
    const T *sB = stage + lenA;
    int i = (int)merge_path(sA, (int64_t)lenA, sB, (int64_t)lenB, (int64_t)o);
    int j = o - i;
    cnt = (cnt_total - o) < kItems ? (cnt_total - o) : kItems;
    merge_seq(sA, (int64_t)lenA, sB, (int64_t)lenB, i, j, tmp, cnt);
  }
  __syncthreads();  // every thread has read its inputs from `stage`
  for (int q = 0; q < cnt; ++q) stage[o + q] = tmp[q];
  __syncthreads();
  for (int t = threadIdx.x; t < cnt_total; t += kSortBlock)
    out[o0 + t] = stage[t];
}


}  // namespace

template <typename T>
void launch_merge_sort(int64_t n, const T *in, T *out, hipStream_t s) {
  if (n <= 0) return;
  const int64_t tiles = ceil_div<int64_t>(n, kTile);
  int passes = 0;
  for (int64_t L = kTile; L < n; L <<= 1) ++passes;
  if (passes == 0) {
    hipLaunchKernelGGL(block_sort_krn<T>, dim3(tiles), dim3(kSortBlock), 0, s,
                       n, in, out);
    KMD_LAUNCH_CHECK();
    return;
  }
  TempBuffer tmp(sizeof(T) * n, s);
  // choose the first destination so that the last pass lands in `out`
  T *bufs[2] = {out, tmp.as<T>()};
  int cur = (passes % 2 == 0) ? 0 : 1;
  hipLaunchKernelGGL(block_sort_krn<T>, dim3(tiles), dim3(kSortBlock), 0, s, n,
                     in, bufs[cur]);
  KMD_LAUNCH_CHECK();
  for (int64_t L = kTile; L < n; L <<= 1) {
    hipLaunchKernelGGL(merge_pass_krn<T>, dim3(tiles), dim3(kSortBlock), 0, s,
                       n, L, (const T *)bufs[cur], bufs[cur ^ 1]);
    KMD_LAUNCH_CHECK();
    cur ^= 1;
  }
}

// ---- order statistic by radix select ---------------------------------------
// The element at position p of the stable ascending sort, without sorting:
// an MSB-first radix select over the sort's order keys (8-bit digits) finds
// its key K and its rank r among the elements whose key is K; the r-th such
// element in input order is the one the stable sort puts at p (elements with
// equal keys differ only as -0 / +0 or NaN payloads, and keep input order).
constexpr int kSelBlock = 256;
constexpr int kSelChunk = 16 * kSelBlock;   // elements per block in the scans

template <typename T>
struct SelState {
  typename Key<T>::U prefix, mask;
  int64_t rank;                       // rank of the target among the prefix
  unsigned long long hist[256];
};

template <typename T>
__global__ void __launch_bounds__(kSelBlock) sel_init_krn(SelState<T> *st, int64_t p) {
  st->hist[threadIdx.x] = 0ull;
  if (threadIdx.x == 0) {
    st->prefix = 0;
    st->mask = 0;
    st->rank = p;
  }
}

template <typename T>
__global__ void __launch_bounds__(kSelBlock)
    sel_hist_krn(int64_t n, const T *__restrict__ data, SelState<T> *st, int shift) {
  using U = typename Key<T>::U;
  __shared__ unsigned h[256];
  h[threadIdx.x] = 0u;
  __syncthreads();
  const U prefix = st->prefix, mask = st->mask;
  // runs of one digit are counted in a register and added once (the first
  // digit is nearly constant over distances, and per-element LDS atomics on
  // one bin would serialise)
  unsigned cur = 0u, run = 0u;
  for (int64_t i = blockIdx.x * (int64_t)kSelBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kSelBlock) {
    const U k = Key<T>::of(data[i]);
    if ((k & mask) == prefix) {
      const unsigned dg = (unsigned)(k >> shift) & 255u;
      if (dg == cur) {
        ++run;
      } else {
        if (run) atomicAdd(&h[cur], run);
        cur = dg;
        run = 1u;
      }
    }
  }
  if (run) atomicAdd(&h[cur], run);
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&st->hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

// one block: the digit whose bucket holds the target; clears the histogram
template <typename T>
__global__ void __launch_bounds__(kSelBlock) sel_pick_krn(SelState<T> *st, int shift) {
  using U = typename Key<T>::U;
  __shared__ unsigned long long h[256];
  h[threadIdx.x] = st->hist[threadIdx.x];
  st->hist[threadIdx.x] = 0ull;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t r = st->rank;
    int b = 0;
    for (; b < 255; ++b) {
      if (r < (int64_t)h[b]) break;
      r -= (int64_t)h[b];
    }
    st->prefix |= (U)b << shift;
    st->mask |= (U)255 << shift;
    st->rank = r;
  }
}

// per chunk: elements whose key is the selected key
template <typename T>
__global__ void __launch_bounds__(kSelBlock)
    sel_count_krn(int64_t n, const T *__restrict__ data, const SelState<T> *st,
                  unsigned *counts) {
  using U = typename Key<T>::U;
  __shared__ unsigned c;
  if (threadIdx.x == 0) c = 0u;
  __syncthreads();
  const U K = st->prefix;
  const int64_t base = (int64_t)blockIdx.x * kSelChunk;
  unsigned mine = 0u;
  for (int j = threadIdx.x; j < kSelChunk; j += kSelBlock) {
    const int64_t i = base + j;
    mine += (i < n && Key<T>::of(data[i]) == K) ? 1u : 0u;
  }
  if (mine) atomicAdd(&c, mine);
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

// one block: the chunk holding the rank-th match, then the match itself
template <typename T>
__global__ void __launch_bounds__(kSelBlock)
    sel_finish_krn(int64_t n, int64_t n_chunks, const T *__restrict__ data,
                   const SelState<T> *st, const unsigned *counts, T *out) {
  using U = typename Key<T>::U;
  __shared__ int64_t s_chunk, s_r;
  __shared__ unsigned long long part[kSelBlock];
  // counts of each thread's contiguous range of chunks, scanned in order
  const int64_t per = ceil_div<int64_t>(n_chunks, kSelBlock);
  const int64_t lo = std::min<int64_t>((int64_t)threadIdx.x * per, n_chunks);
  const int64_t hi = std::min<int64_t>(lo + per, n_chunks);
  unsigned long long mine = 0ull;
  for (int64_t b = lo; b < hi; ++b) mine += counts[b];
  part[threadIdx.x] = mine;
  if (threadIdx.x == 0) s_chunk = -1;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t r = st->rank;
    int t = 0;
    for (; t < kSelBlock - 1; ++t) {
      if (r < (int64_t)part[t]) break;
      r -= (int64_t)part[t];
    }
    s_r = r;
    s_chunk = t;                      // the thread whose range holds it
  }
  __syncthreads();
  if ((int)threadIdx.x == s_chunk) {
    int64_t r = s_r, b = lo;
    for (; b < hi - 1; ++b) {
      if (r < (int64_t)counts[b]) break;
      r -= (int64_t)counts[b];
    }
    s_r = r;
    s_chunk = b;
  }
  __syncthreads();
  const U K = st->prefix;
  const int64_t base = s_chunk * kSelChunk;
  const int lane = threadIdx.x & (kWave - 1);
  if (threadIdx.x < kWave) {          // one wave walks the chunk in order
    int64_t r = s_r;
    for (int j0 = 0; j0 < kSelChunk; j0 += kWave) {
      const int64_t i = base + j0 + lane;
      const bool m = i < n && Key<T>::of(data[i]) == K;
      const unsigned long long bal = __ballot(m);
      const int cnt = __popcll(bal);
      if (r < cnt) {
        // the (r+1)-th set bit of bal
        unsigned long long b = bal;
        for (int q = 0; q < r; ++q) b &= b - 1ull;
        const int pos = __ffsll((long long)b) - 1;
        if (lane == pos) out[0] = data[i];
        break;
      }
      r -= cnt;
    }
  }
}

template <typename T>
void launch_order_statistic(int64_t n, const T *data, int64_t p, T *out,
                            hipStream_t s) {
  using U = typename Key<T>::U;
  constexpr int kBits = 8 * (int)sizeof(U);
  TempBuffer st(sizeof(SelState<T>), s);
  hipLaunchKernelGGL(sel_init_krn<T>, dim3(1), dim3(kSelBlock), 0, s,
                     st.as<SelState<T>>(), p);
  KMD_LAUNCH_CHECK();
  const int64_t want = ceil_div<int64_t>(n, kSelBlock * 16);
  const unsigned grid = (unsigned)std::min<int64_t>(
      std::max<int64_t>(want, 1), 8 * (int64_t)device_cu_count());
  for (int shift = kBits - 8; shift >= 0; shift -= 8) {
    hipLaunchKernelGGL(sel_hist_krn<T>, dim3(grid), dim3(kSelBlock), 0, s, n,
                       data, st.as<SelState<T>>(), shift);
    KMD_LAUNCH_CHECK();
    hipLaunchKernelGGL(sel_pick_krn<T>, dim3(1), dim3(kSelBlock), 0, s,
                       st.as<SelState<T>>(), shift);
    KMD_LAUNCH_CHECK();
  }
  const int64_t n_chunks = ceil_div<int64_t>(n, kSelChunk);
  require(n_chunks <= INT32_MAX, "too many elements");
  TempBuffer counts(sizeof(unsigned) * (size_t)n_chunks, s);
  hipLaunchKernelGGL(sel_count_krn<T>, dim3((unsigned)n_chunks), dim3(kSelBlock),
                     0, s, n, data, st.as<const SelState<T>>(),
                     counts.as<unsigned>());
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(sel_finish_krn<T>, dim3(1), dim3(kSelBlock), 0, s, n,
                     n_chunks, data, st.as<const SelState<T>>(),
                     counts.as<const unsigned>(), out);
  KMD_LAUNCH_CHECK();
}

// sorted_asc[n - topk] (src/util_kernels.hpp:163-206): the maximum for
// topk == 1, else the order statistic by radix select (the same element the
// stable sort would put there).
template <typename T>
void launch_threshold(int64_t n, const T *data, int64_t topk, T *threshold,
                      hipStream_t s) {
  if (topk <= 0) return;
  if (topk == 1) {
    launch_max<T>(n, data, threshold, s);
    return;
  }
  launch_order_statistic<T>(n, data, n - topk, threshold, s);
}

template void launch_merge_sort<float>(int64_t, const float *, float *,
                                       hipStream_t);
template void launch_merge_sort<double>(int64_t, const double *, double *,
                                        hipStream_t);
template void launch_threshold<float>(int64_t, const float *, int64_t, float *,
                                      hipStream_t);
template void launch_threshold<double>(int64_t, const double *, int64_t,
                                       double *, hipStream_t);

}  // namespace kmd
