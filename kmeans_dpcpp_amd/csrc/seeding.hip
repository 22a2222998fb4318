// Data preparation and greedy k-means++ seeding for the estimator front-end
// (kmeans_dpcpp_amd/estimator.py), the caller of kmeans_lloyd_driver.  The
// reference has no counterpart: its driver takes the initial centroids from
// the caller (python_api/_kmeans_lloyd.cpp:1437).
//
//  * column_stats: per-feature mean and variance (FP64, two passes: mean,
//    then the mean squared deviation) of a row-major sample matrix X (n, d).
//  * transpose_shift: dst (cols, rows) = src (rows, cols)^T - shift[col],
//    through 64 x 64 LDS tiles (both sides coalesced).  Turns sample-major
//    X into the feature-major X_t the Lloyd kernels read, centred.
//  * gather_columns: out_t[:, j] = X_t[:, idx[j]].
//  * k-means++: the greedy k-means++ of scikit-learn 1.7.2
//    (sklearn/cluster/_kmeans.py:174-277).  The first center is drawn with
//    probability proportional to w (numpy's choice: first i with
//    cdf_i > u); each later center draws T candidates with probability
//    proportional to w D^2 (first i with cumsum_i >= u pot, clipped to
//    n - 1), D being the distance to the closest chosen center, and keeps
//    the candidate whose potential sum_i w_i min(D_i^2, |x_i - cand|^2) is
//    smallest (first on ties).  The uniforms come from the caller, so the
//    caller's random stream fixes the centers.  Distances, potentials and
//    prefix sums are FP64.
//
// k-means++ layout.  Samples are cut into segments of kSeg; one launch per
// center: every block scores one segment against the T candidates, writes
// m_t(i) = min(D_i^2, |x_i - cand_t|^2) (FP64, row t of a T x n buffer) and
// the segment sums of w m_t; the last block to finish (atomic ticket) sums
// the potentials, picks the winner t*, copies the center and searches the
// prefix sums of w m_t* (segment level, then inside one segment) for the
// next center's candidates.  Row t* of the buffer is the next launch's D^2,
// so a center costs one pass: read X_t, w and D^2, write T rows.
#include "common.hpp"

namespace kmd {

namespace {

constexpr int kNT = 256;             // threads per block
constexpr int kSeg = 4 * kNT;        // samples per segment (4 per thread)
constexpr int kFC = 256;             // candidate features staged per LDS chunk
constexpr int kMaxTrials = 16;

__device__ inline int lane_of() { return threadIdx.x & (kWave - 1); }

__device__ inline double block_sum_d(double v, double *lds /*[kNT / kWave]*/) {
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  __syncthreads();
  if (lane_of() == 0) lds[threadIdx.x / kWave] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < kNT / kWave; ++i) t += lds[i];
  return t;
}

// Loads written by other blocks of the same launch, read after the ticket.
__device__ inline double load_agent(const double *p) {
  return __longlong_as_double((long long)__hip_atomic_load(
      reinterpret_cast<const unsigned long long *>(p), __ATOMIC_RELAXED,
      __HIP_MEMORY_SCOPE_AGENT));
}

// ---- per-feature mean and variance -----------------------------------------
// Block (x, y) covers features [y FC, y FC + FC) (FC = min(d, 256)) of rows
// x P + r, r < P = 256 / FC, striding over rows by gridDim.x P.
template <typename T>
__global__ void __launch_bounds__(kNT)
    col_moment_krn(int64_t n, int64_t d, const T *X, const double *mean,
                   double *part) {
  __shared__ double red[kNT];
  const int fc = (int)(d < kNT ? d : kNT);
  const int P = kNT / fc;
  const int r = threadIdx.x / fc, fl = threadIdx.x % fc;
  const int64_t f = (int64_t)blockIdx.y * fc + fl;
  double acc = 0.0;
  if (r < P && f < d) {
    const double mu = mean ? mean[f] : 0.0;
    for (int64_t row = (int64_t)blockIdx.x * P + r; row < n;
         row += (int64_t)gridDim.x * P) {
      const double x = (double)X[row * d + f];
      if (mean) {
        const double dx = x - mu;
        acc = fma(dx, dx, acc);
      } else {
        acc += x;
      }
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  if (r == 0 && f < d) {
    double s = acc;
    for (int q = 1; q < P; ++q) s += red[q * fc + fl];
    part[(int64_t)blockIdx.x * d + f] = s;
  }
}

// out[f] = sum over blocks of part[b][f] / n
__global__ void __launch_bounds__(kNT)
    col_finish_krn(int64_t n, int64_t d, int64_t nblk, const double *part,
                   double *out) {
  const int64_t f = (int64_t)blockIdx.x * kNT + threadIdx.x;
  if (f >= d) return;
  double s = 0.0;
  for (int64_t b = 0; b < nblk; ++b) s += part[b * d + f];
  out[f] = n > 0 ? s / (double)n : 0.0;
}

// ---- transpose with per-column shift ----------------------------------------
template <typename T>
__global__ void __launch_bounds__(kNT)
    transpose_shift_krn(int64_t rows, int64_t cols, const T *src,
                        const double *shift, T *dst) {
  __shared__ T tile[64][65];
  const int64_t tiles_c = ceil_div<int64_t>(cols, 64);
  const int64_t tr = blockIdx.x / tiles_c, tc = blockIdx.x % tiles_c;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
  const int64_t c = tc * 64 + tx;
  const T sh = (shift && c < cols) ? (T)shift[c] : (T)0;
  for (int j = ty; j < 64; j += 4) {
    const int64_t r = tr * 64 + j;
    if (r < rows && c < cols) tile[j][tx] = src[r * cols + c] - sh;
  }
  __syncthreads();
  const int64_t r_out = tr * 64 + tx;
  for (int j = ty; j < 64; j += 4) {
    const int64_t c_out = tc * 64 + j;
    if (r_out < rows && c_out < cols) dst[c_out * rows + r_out] = tile[tx][j];
  }
}

// ---- column gather -----------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kNT)
    gather_cols_krn(int64_t n, int64_t d, int64_t m, const T *X_t,
                    const int64_t *idx, T *out_t) {
  const int64_t e = (int64_t)blockIdx.x * kNT + threadIdx.x;
  if (e >= d * m) return;
  const int64_t f = e / m, j = e % m;
  out_t[e] = X_t[f * n + idx[j]];
}

// ---- k-means++ ----------------------------------------------------------------
struct KppArgs {
  int64_t n, d, k, nb;
  int trials;                 // T (candidates per center after the first)
  const void *X_t, *w;
  const double *u;            // uniforms: [0] first center, 1 + (c-1) T + j
  void *centers_t;            // (d, k)
  int64_t *indices;           // (k)
  double *m[2];               // [T][n] each: m_t(i) of even / odd centers
  double *part;               // [T][nb] segment sums of w m_t
  int64_t *cand;              // [T] candidates of the current center
  int *tstar;                 // winner of the previous center
  unsigned *ticket;
};

// Per-block search of the prefix sums of v (segment sums seg[nb]; values
// v_i = w_i, or w_i m_i when m != null) for nt targets r[j]: the first i with
// P(i) > r (right) or P(i) >= r (left), else n - 1.  P(i) = (prefix of the
// segment sums up to i's segment, as this block computes it) + (prefix of v
// inside the segment).
template <typename T>
__device__ void block_search(int64_t n, int64_t nb, const double *seg,
                             const T *w, const double *m, const double *r,
                             int nt, bool right, int64_t *out,
                             double *s_chunk /*[kNT]*/, int *s_seg /*[kMaxTrials]*/,
                             double *s_before /*[kMaxTrials]*/) {
  const int tid = threadIdx.x;
  const int64_t per = ceil_div<int64_t>(nb, kNT);
  const int64_t lo = tid * per, hi = (lo + per < nb) ? lo + per : nb;
  double tsum = 0.0;
  for (int64_t s = lo; s < hi; ++s) tsum += load_agent(seg + s);
  if (tid < nt) s_seg[tid] = INT32_MAX;
  s_chunk[tid] = tsum;
  __syncthreads();
  if (tid == 0) {  // exclusive scan of the chunk sums, in chunk order
    double run = 0.0;
    for (int t = 0; t < kNT; ++t) {
      const double v = s_chunk[t];
      s_chunk[t] = run;
      run += v;
    }
  }
  __syncthreads();
  // first segment whose walked prefix reaches r[j]: every thread walks its
  // chunk (from the scanned chunk offset); the lowest hit wins, and its
  // thread records the prefix before that segment
  for (int j = 0; j < nt; ++j) {
    const double rj = r[j];
    int fnd = INT32_MAX;
    double bef = 0.0, run = s_chunk[tid];
    for (int64_t s = lo; s < hi; ++s) {
      const double nxt = run + load_agent(seg + s);
      if (right ? (nxt > rj) : (nxt >= rj)) {
        fnd = (int)s;
        bef = run;
        break;
      }
      run = nxt;
    }
    if (fnd != INT32_MAX) atomicMin(&s_seg[j], fnd);
    __syncthreads();
    if (fnd != INT32_MAX && fnd == s_seg[j]) s_before[j] = bef;
  }
  __syncthreads();
  // inside the segment: wave j % 4 takes target j, 16 samples per lane
  const int wv = tid / kWave, lane = lane_of();
  constexpr int kPer = kSeg / kWave;
  for (int j = wv; j < nt; j += kNT / kWave) {
    const int s = s_seg[j];
    if (s == INT32_MAX) {
      if (lane == 0) out[j] = n - 1;
      continue;
    }
    const int64_t base = (int64_t)s * kSeg;
    const int64_t end = (base + kSeg < n) ? base + kSeg : n;
    const int64_t i0 = base + (int64_t)lane * kPer;
    double v[kPer];
    double lsum = 0.0;
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
      const int64_t i = i0 + e;
      double x = 0.0;
      if (i < end) {
        x = (double)w[i];
        if (m) x *= load_agent(m + i);
      }
      v[e] = x;
      lsum += x;
    }
    // inclusive scan of the lane sums
    double inc = lsum;
    for (int o = 1; o < kWave; o <<= 1) {
      const double y = __shfl_up(inc, o, kWave);
      if (lane >= o) inc += y;
    }
    double q = __shfl_up(inc, 1, kWave);
    if (lane == 0) q = 0.0;
    q += s_before[j];
    int hit = -1;
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
      q += v[e];
      const bool h = (i0 + e < end) && (right ? (q > r[j]) : (q >= r[j]));
      if (h && hit < 0) hit = e;
    }
    const unsigned long long mask = __ballot(hit >= 0);
    if (mask == 0ull) {
      if (lane == 0) out[j] = end - 1;
    } else {
      const int first = __ffsll((long long)mask) - 1;
      if (lane == first) out[j] = i0 + hit;
    }
  }
}

// Segment sums of w, then (last block) the first center.
template <typename T>
__global__ void __launch_bounds__(kNT) kpp_first_krn(KppArgs a) {
  __shared__ double lds[kNT / kWave];
  __shared__ double s_chunk[kNT];
  __shared__ int s_seg[kMaxTrials];
  __shared__ double s_before[kMaxTrials];
  __shared__ double s_r[1];
  __shared__ bool last;
  const T *w = static_cast<const T *>(a.w);
  const int64_t i0 = (int64_t)blockIdx.x * kSeg + 4 * threadIdx.x;
  double v = 0.0;
  for (int j = 0; j < 4; ++j)
    if (i0 + j < a.n) v += (double)w[i0 + j];
  const double s = block_sum_d(v, lds);
  if (threadIdx.x == 0) {
    a.part[blockIdx.x] = s;
    __threadfence();
    last = atomicAdd(a.ticket, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  double t = 0.0;
  for (int64_t b = threadIdx.x; b < a.nb; b += kNT) t += load_agent(a.part + b);
  const double tot = block_sum_d(t, lds);
  if (threadIdx.x == 0) s_r[0] = a.u[0] * tot;
  __syncthreads();
  block_search<T>(a.n, a.nb, a.part, w, nullptr, s_r, 1, true, a.cand, s_chunk,
                  s_seg, s_before);
  if (threadIdx.x == 0) a.ticket[0] = 0u;
}

// One center: scores the segment against the candidates, then (last block)
// picks the winner and the next candidates.
template <typename T, int TT>
__global__ void __launch_bounds__(kNT) kpp_step_krn(KppArgs a, int c) {
  __shared__ double cl[TT * kFC];
  __shared__ double lds[kNT / kWave];
  __shared__ double s_chunk[kNT];
  __shared__ int s_seg[kMaxTrials];
  __shared__ double s_before[kMaxTrials];
  __shared__ double s_pot[kMaxTrials];
  __shared__ double s_r[kMaxTrials];
  __shared__ int64_t s_idx;
  __shared__ int s_ts;
  __shared__ bool last;
  const T *X = static_cast<const T *>(a.X_t);
  const T *w = static_cast<const T *>(a.w);
  const int64_t n = a.n, d = a.d;
  const int nc = (c == 0) ? 1 : a.trials;  // candidates of this center
  const double *dold =
      (c == 0) ? nullptr : a.m[(c - 1) & 1] + (int64_t)a.tstar[0] * n;
  double *mout = a.m[c & 1];
  const int64_t i0 = (int64_t)blockIdx.x * kSeg + 4 * threadIdx.x;
  const bool vec = (i0 + 3 < n) && (n % 4 == 0) &&
                   ((reinterpret_cast<uintptr_t>(X) & 15u) == 0);
  double acc[4][TT];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int t = 0; t < TT; ++t) acc[j][t] = 0.0;

  for (int64_t f0 = 0; f0 < d; f0 += kFC) {
    const int fcn = (int)((d - f0 < kFC) ? d - f0 : kFC);
    __syncthreads();
    for (int e = threadIdx.x; e < TT * fcn; e += kNT) {
      const int t = e / fcn, f = e % fcn;
      cl[t * kFC + f] = (t < nc) ? (double)X[(f0 + f) * n + a.cand[t]] : 0.0;
    }
    __syncthreads();
    for (int f = 0; f < fcn; ++f) {
      const T *row = X + (f0 + f) * n + i0;
      double x[4];
      if constexpr (sizeof(T) == 4) {
        if (vec) {
          const float4 v = *reinterpret_cast<const float4 *>(row);
          x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
        } else {
#pragma unroll
          for (int j = 0; j < 4; ++j) x[j] = (i0 + j < n) ? (double)row[j] : 0.0;
        }
      } else {
        if (vec) {
          const double2 v0 = *reinterpret_cast<const double2 *>(row);
          const double2 v1 = *reinterpret_cast<const double2 *>(row + 2);
          x[0] = v0.x; x[1] = v0.y; x[2] = v1.x; x[3] = v1.y;
        } else {
#pragma unroll
          for (int j = 0; j < 4; ++j) x[j] = (i0 + j < n) ? (double)row[j] : 0.0;
        }
      }
#pragma unroll
      for (int t = 0; t < TT; ++t) {
        if (t < nc) {
          const double cv = cl[t * kFC + f];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const double dx = x[j] - cv;
            acc[j][t] = fma(dx, dx, acc[j][t]);
          }
        }
      }
    }
  }

  // m_t = min(D^2, dist_t); segment sums of w m_t
  double wv[4], dv[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const bool in = i0 + j < n;
    wv[j] = in ? (double)w[i0 + j] : 0.0;
    dv[j] = (in && dold) ? dold[i0 + j] : __builtin_inf();
  }
#pragma unroll
  for (int t = 0; t < TT; ++t) {
    if (t < nc) {
      double mv[4], p = 0.0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double dd = acc[j][t];
        mv[j] = (dv[j] <= dd) ? dv[j] : dd;
        p += (i0 + j < n) ? wv[j] * mv[j] : 0.0;
      }
      double *mo = mout + (int64_t)t * n + i0;
      if (vec) {
        *reinterpret_cast<double2 *>(mo) = make_double2(mv[0], mv[1]);
        *reinterpret_cast<double2 *>(mo + 2) = make_double2(mv[2], mv[3]);
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (i0 + j < n) mo[j] = mv[j];
      }
      const double s = block_sum_d(p, lds);
      if (threadIdx.x == 0) a.part[(int64_t)t * a.nb + blockIdx.x] = s;
    }
  }
  // every thread wrote m rows the last block reads: release from all waves
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(a.ticket, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();

  // potentials, winner, center
  for (int t = 0; t < nc; ++t) {
    double v = 0.0;
    for (int64_t b = threadIdx.x; b < a.nb; b += kNT)
      v += load_agent(a.part + (int64_t)t * a.nb + b);
    const double tot = block_sum_d(v, lds);
    if (threadIdx.x == 0) s_pot[t] = tot;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int ts = 0;
    for (int t = 1; t < nc; ++t)
      if (s_pot[t] < s_pot[ts]) ts = t;
    s_ts = ts;
    s_idx = a.cand[ts];
    a.tstar[0] = ts;
    a.indices[c] = s_idx;
  }
  __syncthreads();
  T *C = static_cast<T *>(a.centers_t);
  for (int64_t f = threadIdx.x; f < d; f += kNT) C[f * a.k + c] = X[f * n + s_idx];
  if (c + 1 < a.k) {
    const int ts = s_ts;
    if (threadIdx.x < a.trials)
      s_r[threadIdx.x] = a.u[1 + (int64_t)c * a.trials + threadIdx.x] * s_pot[ts];
    __syncthreads();
    block_search<T>(n, a.nb, a.part + (int64_t)ts * a.nb, w,
                    mout + (int64_t)ts * n, s_r, a.trials, false, a.cand,
                    s_chunk, s_seg, s_before);
  }
  if (threadIdx.x == 0) a.ticket[0] = 0u;
}

template <typename T, int TT>
void launch_kpp_step(const KppArgs &a, int c, hipStream_t s) {
  hipLaunchKernelGGL((kpp_step_krn<T, TT>), dim3((unsigned)a.nb), dim3(kNT), 0,
                     s, a, c);
  KMD_LAUNCH_CHECK();
}

}  // namespace

// ---- launchers ------------------------------------------------------------------
template <typename T>
void launch_column_mean_var(int64_t n, int64_t d, const T *X, double *mean,
                            double *var, hipStream_t s) {
  if (d <= 0) return;
  const int fc = (int)(d < kNT ? d : kNT);
  const int P = kNT / fc;
  const int64_t gy = ceil_div<int64_t>(d, fc);
  int64_t gx = ceil_div<int64_t>(n > 0 ? n : 1, P);
  const int64_t cap = ceil_div<int64_t>(4 * device_cu_count(), gy);
  if (gx > cap) gx = cap > 0 ? cap : 1;
  TempBuffer part((size_t)(gx * d) * sizeof(double), s);
  const dim3 grid((unsigned)gx, (unsigned)gy);
  const unsigned fb = (unsigned)ceil_div<int64_t>(d, kNT);
  hipLaunchKernelGGL(col_moment_krn<T>, grid, dim3(kNT), 0, s, n, d, X,
                     (const double *)nullptr, part.as<double>());
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(col_finish_krn, dim3(fb), dim3(kNT), 0, s, n, d, gx,
                     (const double *)part.as<double>(), mean);
  KMD_LAUNCH_CHECK();
  if (var) {
    hipLaunchKernelGGL(col_moment_krn<T>, grid, dim3(kNT), 0, s, n, d, X,
                       (const double *)mean, part.as<double>());
    KMD_LAUNCH_CHECK();
    hipLaunchKernelGGL(col_finish_krn, dim3(fb), dim3(kNT), 0, s, n, d, gx,
                       (const double *)part.as<double>(), var);
    KMD_LAUNCH_CHECK();
  }
}

template <typename T>
void launch_transpose_shift(int64_t rows, int64_t cols, const T *src,
                            const double *shift, T *dst, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return;
  const int64_t tiles = ceil_div<int64_t>(rows, 64) * ceil_div<int64_t>(cols, 64);
  require(tiles <= INT32_MAX, "array too large to transpose");
  hipLaunchKernelGGL(transpose_shift_krn<T>, dim3((unsigned)tiles), dim3(kNT),
                     0, s, rows, cols, src, shift, dst);
  KMD_LAUNCH_CHECK();
}

template <typename T>
void launch_gather_columns(int64_t n, int64_t d, int64_t m, const T *X_t,
                           const int64_t *idx, T *out_t, hipStream_t s) {
  if (d <= 0 || m <= 0) return;
  const int64_t blocks = ceil_div<int64_t>(d * m, kNT);
  hipLaunchKernelGGL(gather_cols_krn<T>, dim3((unsigned)blocks), dim3(kNT), 0,
                     s, n, d, m, X_t, idx, out_t);
  KMD_LAUNCH_CHECK();
}

int kpp_max_trials() { return kMaxTrials; }

template <typename T>
void launch_kmeans_plusplus(int64_t n, int64_t d, int64_t k, int trials,
                            const T *X_t, const T *w, const double *u_dev,
                            T *centers_t, int64_t *indices, hipStream_t s) {
  require(trials >= 1 && trials <= kMaxTrials,
          "n_local_trials must be between 1 and 16");
  require(k >= 1 && n >= k, "n_samples must be >= n_clusters >= 1");
  const int64_t nb = ceil_div<int64_t>(n, kSeg);
  require(nb <= INT32_MAX, "too many samples");
  const size_t mrow = (size_t)trials * (size_t)n;
  TempBuffer mbuf(2 * mrow * sizeof(double), s);
  TempBuffer part((size_t)trials * (size_t)nb * sizeof(double), s);
  TempBuffer small(256, s);
  KppArgs a{};
  a.n = n; a.d = d; a.k = k; a.nb = nb; a.trials = trials;
  a.X_t = X_t; a.w = w; a.u = u_dev; a.centers_t = centers_t;
  a.indices = indices;
  a.m[0] = mbuf.as<double>();
  a.m[1] = mbuf.as<double>() + mrow;
  a.part = part.as<double>();
  char *sm = small.as<char>();
  a.cand = reinterpret_cast<int64_t *>(sm);              // 16 x 8 bytes
  a.tstar = reinterpret_cast<int *>(sm + 128);
  a.ticket = reinterpret_cast<unsigned *>(sm + 136);
  KMD_HIP_CHECK(hipMemsetAsync(sm, 0, 256, s));
  hipLaunchKernelGGL(kpp_first_krn<T>, dim3((unsigned)nb), dim3(kNT), 0, s, a);
  KMD_LAUNCH_CHECK();
  launch_kpp_step<T, 1>(a, 0, s);
  for (int64_t c = 1; c < k; ++c) {
    if (trials <= 2)
      launch_kpp_step<T, 2>(a, (int)c, s);
    else if (trials <= 4)
      launch_kpp_step<T, 4>(a, (int)c, s);
    else if (trials <= 8)
      launch_kpp_step<T, 8>(a, (int)c, s);
    else
      launch_kpp_step<T, 16>(a, (int)c, s);
  }
}

#define KMD_SEED_INST(T)                                                       \
  template void launch_column_mean_var<T>(int64_t, int64_t, const T *,         \
                                          double *, double *, hipStream_t);    \
  template void launch_transpose_shift<T>(int64_t, int64_t, const T *,         \
                                          const double *, T *, hipStream_t);   \
  template void launch_gather_columns<T>(int64_t, int64_t, int64_t, const T *, \
                                         const int64_t *, T *, hipStream_t);   \
  template void launch_kmeans_plusplus<T>(int64_t, int64_t, int64_t, int,      \
                                          const T *, const T *, const double *,\
                                          T *, int64_t *, hipStream_t);
KMD_SEED_INST(float)
KMD_SEED_INST(double)
#undef KMD_SEED_INST

}  // namespace kmd
