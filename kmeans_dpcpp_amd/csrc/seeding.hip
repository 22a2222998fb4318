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
// k-means++ layout.  Samples are cut into segments of kSeg; two launches per
// center: a scoring pass where every block scores one segment against the
// T candidates and writes m_t(i) = min(D_i^2, |x_i - cand_t|^2) (FP64, row
// t of a T x n buffer) and the segment sums of w m_t; then a one-block pick
// that sums the potentials, takes the winner t*, copies the center and
// searches the prefix sums of w m_t* (segment level, then inside one
// segment) for the next center's candidates.  Row t* of the buffer is the
// next center's D^2, so a center costs one pass over the data: read X_t, w
// and D^2, write T rows.  (The kernel boundary publishes the rows; a
// last-block ticket inside one launch would need a device-scope release in
// every block, which stalls each CU on an L2 write-back.)
#include "common.hpp"

namespace kmd {

namespace {

constexpr int kNT = 256;             // threads per block (data preparation)
constexpr int kFC = 64;              // candidate features staged per LDS chunk
constexpr int kMaxTrials = 16;

__device__ inline int lane_of() { return threadIdx.x & (kWave - 1); }

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
constexpr int kSNT = 512;                 // threads per scoring block (2 per CU)
constexpr int kPNT = 1024;                // threads of the one-block picks
constexpr int kSPT = 4;                   // samples per scoring thread
constexpr int kSeg = kSPT * kSNT;         // samples per segment (2048)
constexpr int kSub = 64;                  // samples per sub-segment
constexpr int kSubPerSeg = kSeg / kSub;   // 32
constexpr int kUF = 4;                    // feature rows loaded together
constexpr int kStage = 4;                 // rows per LDS DMA stage (float)
constexpr int kPickD = 1024;              // center features the pick keeps in LDS

static_assert(kSubPerSeg <= kWave, "one lane per sub-segment of a segment");
static_assert(kSub == kWave, "one lane per sample of a sub-segment");

template <typename T>
struct Quad {
  T v[kSPT];
};

// Samples i0 .. i0 + 3 of one feature row (zero past n).
template <typename T>
__device__ inline Quad<T> load_quad(const T *row, int64_t i0, int64_t n, bool vec) {
  Quad<T> q;
  if (vec) {
    if constexpr (sizeof(T) == 4) {
      const float4 v = *reinterpret_cast<const float4 *>(row);
      q.v[0] = v.x; q.v[1] = v.y; q.v[2] = v.z; q.v[3] = v.w;
    } else {
      const double2 v0 = *reinterpret_cast<const double2 *>(row);
      const double2 v1 = *reinterpret_cast<const double2 *>(row + 2);
      q.v[0] = v0.x; q.v[1] = v0.y; q.v[2] = v1.x; q.v[3] = v1.y;
    }
  } else {
#pragma unroll
    for (int j = 0; j < kSPT; ++j) q.v[j] = (i0 + j < n) ? row[j] : (T)0;
  }
  return q;
}

struct KppArgs {
  int64_t n, d, k, nb, nsub;
  int trials;                 // T (candidates per center after the first)
  const void *X_t, *w;
  const double *u;            // uniforms: [0] first center, 1 + (c-1) T + j
  void *centers_t;            // (d, k)
  int64_t *indices;           // (k)
  double *d2[2];              // D^2 to centers 0..c-1, written by pass c into d2[c & 1]
  double *part;               // [T][nb] segment sums of w m_t
  double *part2;              // [T][nsub] sub-segment sums of w m_t
  int64_t *cand;              // [T] candidates of the current center
  double *csq;                // [T + 1] |staged row|^2 of the last scoring pass
};

struct SearchLds {
  double wtot[kPNT / kWave];  // wave totals, then their exclusive prefix
  double r[kMaxTrials];
  double before[kMaxTrials];
  int seg[kMaxTrials];
};

__device__ inline double wave_incl_scan(double y, int lane) {
  for (int o = 1; o < kWave; o <<= 1) {
    const double t = __shfl_up(y, o, kWave);
    if (lane >= o) y += t;
  }
  return y;
}

// Search of the prefix sums of v (value(i) = w_i or w_i m_i) for nt targets
// L.r[j]: the first i with P(i) > r (right) or P(i) >= r (left), else n - 1.
// Three levels, each defined by this search's own arithmetic: the segment
// sums seg[nb], each thread owning a contiguous range (block-scanned range
// totals, then a sequential walk), the first segment whose inclusive prefix
// reaches r; inside it the sub-segment sums sub[nsub] (32 per segment), the
// first reaching r from the segment's exclusive prefix (else its last);
// inside that the 64 values, the first reaching r (else the last).  One
// block; writes out[0 : nt].
template <typename VF>
__device__ void pick_search(int64_t n, int64_t nb, int64_t nsub, const double *seg,
                            const double *sub, VF value, int nt, bool right,
                            int64_t *out, SearchLds &L) {
  const int tid = threadIdx.x, lane = lane_of(), wv = tid / kWave;
  const int64_t per = ceil_div<int64_t>(nb, kPNT);
  const int64_t lo = ((int64_t)tid * per < nb) ? (int64_t)tid * per : nb;
  const int64_t hi = (lo + per < nb) ? lo + per : nb;
  constexpr int kKeep = 8;              // range values kept in registers
  double keep[kKeep];
  double tsum = 0.0;
#pragma unroll
  for (int q = 0; q < kKeep; ++q) {
    keep[q] = (lo + q < hi) ? seg[lo + q] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < kKeep; ++q)
    if (lo + q < hi) tsum += keep[q];
  for (int64_t s = lo + kKeep; s < hi; ++s) tsum += seg[s];
  const double y = wave_incl_scan(tsum, lane);
  double yp = __shfl_up(y, 1, kWave);
  if (lane == 0) yp = 0.0;
  if (lane == kWave - 1) L.wtot[wv] = y;
  if (tid < nt) L.seg[tid] = INT32_MAX;
  __syncthreads();
  if (tid == 0) {                                   // exclusive wave prefix
    double run = 0.0;
    for (int i = 0; i < kPNT / kWave; ++i) {
      const double t = L.wtot[i];
      L.wtot[i] = run;
      run += t;
    }
  }
  __syncthreads();
  int found[kMaxTrials];
  double bef[kMaxTrials], rj[kMaxTrials];
#pragma unroll
  for (int j = 0; j < kMaxTrials; ++j) {
    found[j] = INT32_MAX;
    bef[j] = 0.0;
    rj[j] = (j < nt) ? L.r[j] : 0.0;
  }
  double run = L.wtot[wv] + yp;                     // prefix before this range
  for (int64_t s = lo; s < hi; ++s) {
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < kKeep; ++q)
      if (s == lo + q) v = keep[q];
    const double nxt = run + ((s < lo + kKeep) ? v : seg[s]);
#pragma unroll
    for (int j = 0; j < kMaxTrials; ++j) {
      const bool hit = right ? (nxt > rj[j]) : (nxt >= rj[j]);
      if (j < nt && hit && found[j] == INT32_MAX) {
        found[j] = (int)s;
        bef[j] = run;
      }
    }
    run = nxt;
  }
#pragma unroll
  for (int j = 0; j < kMaxTrials; ++j)
    if (j < nt && found[j] != INT32_MAX) atomicMin(&L.seg[j], found[j]);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kMaxTrials; ++j)
    if (j < nt && found[j] != INT32_MAX && found[j] == L.seg[j]) L.before[j] = bef[j];
  __syncthreads();
  for (int j = wv; j < nt; j += kPNT / kWave) {    // one wave per target
    const int sj = L.seg[j];
    if (sj == INT32_MAX) {
      if (lane == 0) out[j] = n - 1;
      continue;
    }
    const double r = L.r[j];
    // sub-segments of segment sj
    const int64_t b0 = (int64_t)sj * kSubPerSeg;
    const int64_t b_end = (b0 + kSubPerSeg < nsub) ? b0 + kSubPerSeg : nsub;
    const bool bin = (lane < kSubPerSeg) && (b0 + lane < b_end);
    const double yb = wave_incl_scan(bin ? sub[b0 + lane] : 0.0, lane);
    double ybp = __shfl_up(yb, 1, kWave);
    if (lane == 0) ybp = 0.0;
    const double ib = L.before[j] + yb;
    unsigned long long mask = __ballot(bin && (right ? (ib > r) : (ib >= r)));
    const int lb = mask ? __ffsll((long long)mask) - 1 : (int)(b_end - b0 - 1);
    const int64_t bsel = b0 + lb;
    const double before2 = L.before[j] + __shfl(ybp, lb, kWave);
    // samples of sub-segment bsel
    const int64_t i = bsel * kSub + lane;
    const int64_t i_end = (bsel * kSub + kSub < n) ? bsel * kSub + kSub : n;
    const double vi = (i < i_end) ? value(i) : 0.0;
    const double qi = before2 + wave_incl_scan(vi, lane);
    mask = __ballot(i < i_end && (right ? (qi > r) : (qi >= r)));
    if (lane == 0) out[j] = mask ? bsel * kSub + (__ffsll((long long)mask) - 1) : i_end - 1;
  }
}

// Segment and sub-segment sums of w.
template <typename T>
__global__ void __launch_bounds__(kSNT) kpp_wsum_krn(KppArgs a) {
  __shared__ double lds[kSNT / kWave];
  const T *w = static_cast<const T *>(a.w);
  const int64_t i0 = (int64_t)blockIdx.x * kSeg + kSPT * threadIdx.x;
  double v = 0.0;
  for (int j = 0; j < kSPT; ++j)
    if (i0 + j < a.n) v += (double)w[i0 + j];
  for (int o = 1; o < kSub / kSPT; o <<= 1) v += __shfl_xor(v, o, kWave);
  const int64_t sb = i0 / kSub;
  if ((threadIdx.x & (kSub / kSPT - 1)) == 0 && sb < a.nsub) a.part2[sb] = v;
  for (int o = kSub / kSPT; o < kWave; o <<= 1) v += __shfl_xor(v, o, kWave);
  if (lane_of() == 0) lds[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kSNT / kWave; ++i) t += lds[i];
    a.part[blockIdx.x] = t;
  }
}

// One block: the first center, drawn with probability proportional to w.
template <typename T>
__global__ void __launch_bounds__(kPNT) kpp_first_pick_krn(KppArgs a) {
  __shared__ double lds[kPNT / kWave];
  __shared__ SearchLds L;
  const T *w = static_cast<const T *>(a.w);
  double t = 0.0;
  for (int64_t b = threadIdx.x; b < a.nb; b += kPNT) t += a.part[b];
  for (int o = kWave / 2; o > 0; o >>= 1) t += __shfl_xor(t, o, kWave);
  if (lane_of() == 0) lds[threadIdx.x / kWave] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int i = 0; i < kPNT / kWave; ++i) tot += lds[i];
    L.r[0] = a.u[0] * tot;
  }
  pick_search(a.n, a.nb, a.nsub, a.part, a.part2,
              [&](int64_t i) { return (double)w[i]; }, 1, true, a.cand, L);
}

// Squared distance in the scoring pass's arithmetic: |x|^2 + |c|^2 - 2 x.c
// in FP64 with the three sums accumulated by fused multiply-adds over
// features in ascending order, clamped at 0.
__device__ inline double kpp_dist(double xsq, double csq, double xc) {
  const double dd = fma(-2.0, xc, xsq + csq);
  return dd > 0.0 ? dd : 0.0;
}

// One center, scoring pass (pure streaming; the pick is the next launch).
// cache-policy operand of the LDS-DMA row loads: non-temporal (X is read
// once per center and is larger than the caches; D^2 is what is reused)
constexpr int kLoadNT = 2;

// Row 0 of the staged points is center c - 1 (the first center when
// c == 0): D^2 <- min(D^2, dist0) is written to d2[c & 1].  Rows 1..T are
// the candidates of center c: m_t = min(D^2, dist_t) gives the segment and
// sub-segment sums of w m_t (for c == 0, of w D^2 itself).
template <typename T, int TT, bool DMA>
__global__ void __launch_bounds__(kSNT, 4) kpp_score_krn(KppArgs a, int c, int64_t blk0) {
  __shared__ double cl[(TT + 1) * kFC];
  __shared__ double csq[TT + 1];
  // DMA (float, whole aligned segments only): rows staged in LDS,
  // [2 buffers][kStage rows][kSeg samples]; each wave fills and reads only
  // its own 256-sample slice
  __shared__ float xs[DMA ? 2 * kStage * kSeg : 1];
  const int64_t blk = blk0 + blockIdx.x;      // segment of this block
  const T *X = static_cast<const T *>(a.X_t);
  const T *w = static_cast<const T *>(a.w);
  const int64_t n = a.n, d = a.d;
  const int nc = (c == 0) ? 0 : a.trials;   // candidates of this center
  const int rows = nc + 1;
  const int64_t idx0 = (c == 0) ? a.cand[0] : a.indices[c - 1];
  const double *dold = (c == 0) ? nullptr : a.d2[(c - 1) & 1];
  double *dnew = a.d2[c & 1];
  const int64_t i0 = blk * kSeg + kSPT * threadIdx.x;
  const bool vec = (i0 + kSPT - 1 < n) && (n % kSPT == 0) &&
                   ((reinterpret_cast<uintptr_t>(X) & 15u) == 0);
  double acc0[kSPT], acc[kSPT][TT], xsq[kSPT];
#pragma unroll
  for (int j = 0; j < kSPT; ++j) {
    xsq[j] = 0.0;
    acc0[j] = 0.0;
#pragma unroll
    for (int t = 0; t < TT; ++t) acc[j][t] = 0.0;
  }
  if (threadIdx.x <= TT) csq[threadIdx.x] = 0.0;
  auto step = [&](const Quad<T> &xq, int f) {
    double x[kSPT];
#pragma unroll
    for (int j = 0; j < kSPT; ++j) x[j] = (double)xq.v[j];
    const double c0 = cl[f];
#pragma unroll
    for (int j = 0; j < kSPT; ++j) {
      xsq[j] = fma(x[j], x[j], xsq[j]);
      acc0[j] = fma(x[j], c0, acc0[j]);
    }
#pragma unroll
    for (int t = 0; t < TT; ++t) {
      if (t < nc) {
        const double cv = cl[(t + 1) * kFC + f];
#pragma unroll
        for (int j = 0; j < kSPT; ++j) acc[j][t] = fma(x[j], cv, acc[j][t]);
      }
    }
  };

  for (int64_t f0 = 0; f0 < d; f0 += kFC) {
    const int fcn = (int)((d - f0 < kFC) ? d - f0 : kFC);
    __syncthreads();
    for (int e = threadIdx.x; e < (TT + 1) * fcn; e += kSNT) {
      const int r = e / fcn, f = e % fcn;
      const int64_t src = (r == 0) ? idx0 : ((r < rows) ? a.cand[r - 1] : -1);
      cl[r * kFC + f] = (src >= 0) ? (double)X[(f0 + f) * n + src] : 0.0;
    }
    __syncthreads();
    if (threadIdx.x <= TT) {    // |c|^2 of each staged row, features ascending
      double q = csq[threadIdx.x];
      for (int f = 0; f < fcn; ++f) {
        const double cv = cl[threadIdx.x * kFC + f];
        q = fma(cv, cv, q);
      }
      csq[threadIdx.x] = q;
    }
    if constexpr (DMA && sizeof(T) == 4) {
      {
        // rows by LDS DMA (global_load_lds_dwordx4): the next stage of
        // kStage rows is in flight while the current one is scored, with no
        // registers held for it; a wave reads back only what it loaded
        const int wv = threadIdx.x / kWave, ln = lane_of();
        // lanes past n (ragged last segment) copy the row's last 4 samples
        // instead: in bounds and 16-byte aligned (n % 4 == 0); their
        // samples are masked in the epilogue
        const int64_t s_lane = blk * kSeg + wv * (kWave * kSPT) + ln * kSPT;
        const float *gbase = reinterpret_cast<const float *>(X) + f0 * n +
                             (s_lane + kSPT <= n ? s_lane : n - kSPT);
        float *lbase = xs + wv * (kWave * kSPT);
        const int ns = (fcn + kStage - 1) / kStage;
        auto issue = [&](int st) {
          float *buf = lbase + (st & 1) * (kStage * kSeg);
#pragma unroll
          for (int r = 0; r < kStage; ++r) {
            // always kStage copies (a short last stage repeats its last row),
            // so the vmcnt the reads wait for is a constant
            const int f = (st * kStage + r < fcn) ? st * kStage + r : fcn - 1;
            __builtin_amdgcn_global_load_lds(
                (const void *)(gbase + (int64_t)f * n),
                (__attribute__((address_space(3))) void *)(buf + r * kSeg), 16, 0,
                kLoadNT);
          }
        };
        issue(0);
        for (int st = 0; st < ns; ++st) {
          // earlier reads of the buffer the next stage overwrites are done
          asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
          if (st + 1 < ns) {
            issue(st + 1);
            asm volatile("s_waitcnt vmcnt(%0)" ::"n"(kStage) : "memory");
          } else {
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
          }
          const float *buf = xs + (st & 1) * (kStage * kSeg) + threadIdx.x * kSPT;
#pragma unroll
          for (int r = 0; r < kStage; ++r) {
            const int f = st * kStage + r;
            if (f < fcn) {
              const float4 v = *reinterpret_cast<const float4 *>(buf + r * kSeg);
              Quad<T> q;
              q.v[0] = v.x; q.v[1] = v.y; q.v[2] = v.z; q.v[3] = v.w;
              step(q, f);
            }
          }
        }
        continue;
      }
    }
    // groups of UF feature rows, the next group's loads issued before the
    // current group is scored (2 UF rows in flight per lane)
    const T *base = X + f0 * n + i0;
    constexpr int UF0 = (TT <= 4) ? kUF : ((TT <= 7) ? kUF / 2 : 1);
    constexpr int UF = (sizeof(T) == 4 || UF0 == 1) ? UF0 : UF0 / 2;  // registers
    const int fg = fcn - fcn % UF;
    Quad<T> cur[UF], nxt[UF];
    if (fg > 0) {
#pragma unroll
      for (int u = 0; u < UF; ++u) cur[u] = load_quad(base + (int64_t)u * n, i0, n, vec);
    }
    for (int f = 0; f < fg; f += UF) {
      if (f + UF < fg) {
#pragma unroll
        for (int u = 0; u < UF; ++u)
          nxt[u] = load_quad(base + (int64_t)(f + UF + u) * n, i0, n, vec);
      }
#pragma unroll
      for (int u = 0; u < UF; ++u) step(cur[u], f + u);
#pragma unroll
      for (int u = 0; u < UF; ++u) cur[u] = nxt[u];
    }
    for (int f = fg; f < fcn; ++f) step(load_quad(base + (int64_t)f * n, i0, n, vec), f);
  }
  __syncthreads();            // csq complete, cl free
  if (blk == 0 && threadIdx.x <= TT) a.csq[threadIdx.x] = csq[threadIdx.x];

  double wv[kSPT], dv[kSPT], dn[kSPT];
  const bool wvec = vec && ((reinterpret_cast<uintptr_t>(w) & 15u) == 0);
  if (wvec) {                 // vector loads of w and D^2 (aligned like X)
    const Quad<T> wq = load_quad(w + i0, i0, n, true);
#pragma unroll
    for (int j = 0; j < kSPT; ++j) wv[j] = (double)wq.v[j];
    if (dold) {
      const double2 d01 = *reinterpret_cast<const double2 *>(dold + i0);
      const double2 d23 = *reinterpret_cast<const double2 *>(dold + i0 + 2);
      dv[0] = d01.x; dv[1] = d01.y; dv[2] = d23.x; dv[3] = d23.y;
    } else {
#pragma unroll
      for (int j = 0; j < kSPT; ++j) dv[j] = __builtin_inf();
    }
  } else {
#pragma unroll
    for (int j = 0; j < kSPT; ++j) {
      const bool in = i0 + j < n;
      wv[j] = in ? (double)w[i0 + j] : 0.0;
      dv[j] = (in && dold) ? dold[i0 + j] : __builtin_inf();
    }
  }
  const double c0sq = csq[0];
#pragma unroll
  for (int j = 0; j < kSPT; ++j) {
    const double d0 = kpp_dist(xsq[j], c0sq, acc0[j]);
    dn[j] = (dv[j] <= d0) ? dv[j] : d0;
  }
  if (vec) {
    *reinterpret_cast<double2 *>(dnew + i0) = make_double2(dn[0], dn[1]);
    *reinterpret_cast<double2 *>(dnew + i0 + 2) = make_double2(dn[2], dn[3]);
  } else {
#pragma unroll
    for (int j = 0; j < kSPT; ++j)
      if (i0 + j < n) dnew[i0 + j] = dn[j];
  }

  double *wsum = cl;          // [kSNT / kWave][TT]
  const int64_t sb = i0 / kSub;
  const bool sub_lead = (threadIdx.x & (kSub / kSPT - 1)) == 0 && sb < a.nsub;
  const int ns = (c == 0) ? 1 : nc;        // sums to produce
#pragma unroll
  for (int t = 0; t < TT; ++t) {
    if (t < ns) {
      double p = 0.0;
      const double ct = csq[t + 1];
#pragma unroll
      for (int j = 0; j < kSPT; ++j) {
        double mv = dn[j];
        if (c != 0) {
          const double dd = kpp_dist(xsq[j], ct, acc[j][t]);
          mv = (dn[j] <= dd) ? dn[j] : dd;
        }
        p += (i0 + j < n) ? wv[j] * mv : 0.0;
      }
      // sub-segment sum (16 lanes), then the wave sum
      for (int o = 1; o < kSub / kSPT; o <<= 1) p += __shfl_xor(p, o, kWave);
      if (sub_lead) a.part2[(int64_t)t * a.nsub + sb] = p;
      for (int o = kSub / kSPT; o < kWave; o <<= 1) p += __shfl_xor(p, o, kWave);
      if (lane_of() == 0) wsum[(threadIdx.x / kWave) * TT + t] = p;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < ns) {
    double tot = 0.0;
    for (int i = 0; i < kSNT / kWave; ++i) tot += wsum[i * TT + threadIdx.x];
    a.part[(int64_t)threadIdx.x * a.nb + blk] = tot;
  }
}

// One center, pick: potentials, winner (first minimum), center copy, and the
// next center's candidates.  One block.  The values w_i m_t*(i) the last
// search level needs are recomputed for 64 samples per target with the
// scoring pass's arithmetic (the pass does not store m).
template <typename T>
__global__ void __launch_bounds__(kPNT) kpp_pick_krn(KppArgs a, int c) {
  __shared__ double red[kPNT / kWave][kMaxTrials];
  __shared__ SearchLds L;
  __shared__ double s_pot[kMaxTrials];
  __shared__ double s_csq;
  __shared__ int64_t s_idx;
  __shared__ int s_ts;
  __shared__ double s_c[kPickD];               // center c (d <= kPickD)
  const T *X = static_cast<const T *>(a.X_t);
  const T *w = static_cast<const T *>(a.w);
  const int64_t n = a.n, d = a.d, k = a.k;
  const int ns = (c == 0) ? 1 : a.trials;
  // all potentials in one pass: per thread, each trial's segment sums in
  // ascending order; the ns loads of one segment are in flight together
  double v[kMaxTrials];
#pragma unroll
  for (int t = 0; t < kMaxTrials; ++t) v[t] = 0.0;
  for (int64_t b = threadIdx.x; b < a.nb; b += kPNT) {
    double p[kMaxTrials];
#pragma unroll
    for (int t = 0; t < kMaxTrials; ++t) p[t] = (t < ns) ? a.part[(int64_t)t * a.nb + b] : 0.0;
#pragma unroll
    for (int t = 0; t < kMaxTrials; ++t)
      if (t < ns) v[t] += p[t];
  }
#pragma unroll
  for (int t = 0; t < kMaxTrials; ++t) {
    double vt = v[t];
    for (int o = kWave / 2; o > 0; o >>= 1) vt += __shfl_xor(vt, o, kWave);
    if (lane_of() == 0) red[threadIdx.x / kWave][t] = vt;
  }
  __syncthreads();
  if ((int)threadIdx.x < ns) {
    double tot = 0.0;
    for (int i = 0; i < kPNT / kWave; ++i) tot += red[i][threadIdx.x];
    s_pot[threadIdx.x] = tot;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int ts = 0;
    for (int t = 1; t < ns; ++t)
      if (s_pot[t] < s_pot[ts]) ts = t;
    s_ts = ts;
    s_idx = a.cand[ts];
    a.indices[c] = s_idx;
    s_csq = a.csq[ts + 1];                     // |center c|^2 as the pass has it
  }
  __syncthreads();
  const int64_t idx = s_idx;
  T *C = static_cast<T *>(a.centers_t);
  for (int64_t f = threadIdx.x; f < d; f += kPNT) {
    const T v = X[f * n + idx];
    C[f * k + c] = v;
    if (f < kPickD) s_c[f] = (double)v;
  }
  __syncthreads();
  if (c + 1 < k) {
    const int ts = s_ts;
    if ((int)threadIdx.x < a.trials)
      L.r[threadIdx.x] = a.u[1 + (int64_t)c * a.trials + threadIdx.x] * s_pot[ts];
    const double *dn = a.d2[c & 1];
    const double ccsq = s_csq;
    auto value = [&](int64_t i) {
      double m = dn[i];
      if (c != 0) {                            // min(D^2, |x_i - center c|^2)
        double xsq = 0.0, xc = 0.0;
#pragma unroll 16
        for (int64_t f = 0; f < d; ++f) {
          const double x = (double)X[f * n + i];
          const double cf = (f < kPickD) ? s_c[f] : (double)X[f * n + idx];
          xsq = fma(x, x, xsq);
          xc = fma(x, cf, xc);
        }
        const double dd = kpp_dist(xsq, ccsq, xc);
        m = (m <= dd) ? m : dd;
      }
      return (double)w[i] * m;
    };
    pick_search(n, a.nb, a.nsub, a.part + (int64_t)ts * a.nb,
                a.part2 + (int64_t)ts * a.nsub, value, a.trials, false, a.cand, L);
  }
}

template <typename T, int TT>
void launch_kpp_step(const KppArgs &a, int c, hipStream_t s) {
  // 16-byte aligned float rows (n % 4 == 0) stream through LDS DMA; other
  // data through registers
  if (std::is_same<T, float>::value && a.n % kSPT == 0 &&
      (reinterpret_cast<uintptr_t>(a.X_t) & 15u) == 0)
    hipLaunchKernelGGL((kpp_score_krn<T, TT, true>), dim3((unsigned)a.nb),
                       dim3(kSNT), 0, s, a, c, (int64_t)0);
  else
    hipLaunchKernelGGL((kpp_score_krn<T, TT, false>), dim3((unsigned)a.nb),
                       dim3(kSNT), 0, s, a, c, (int64_t)0);
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(kpp_pick_krn<T>, dim3(1), dim3(kPNT), 0, s, a, c);
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
  const int64_t nsub = ceil_div<int64_t>(n, kSub);
  require(nb <= INT32_MAX, "too many samples");
  TempBuffer dbuf(2 * (size_t)n * sizeof(double), s);
  TempBuffer part((size_t)trials * (size_t)(nb + nsub) * sizeof(double), s);
  TempBuffer small(512, s);
  KppArgs a{};
  a.n = n; a.d = d; a.k = k; a.nb = nb; a.nsub = nsub; a.trials = trials;
  a.X_t = X_t; a.w = w; a.u = u_dev; a.centers_t = centers_t;
  a.indices = indices;
  a.d2[0] = dbuf.as<double>();
  a.d2[1] = dbuf.as<double>() + n;
  a.part = part.as<double>();
  a.part2 = part.as<double>() + (size_t)trials * (size_t)nb;
  char *sm = small.as<char>();
  a.cand = reinterpret_cast<int64_t *>(sm);              // 16 x 8 bytes
  a.csq = reinterpret_cast<double *>(sm + 128);          // 17 x 8 bytes
  KMD_HIP_CHECK(hipMemsetAsync(sm, 0, 512, s));
  hipLaunchKernelGGL(kpp_wsum_krn<T>, dim3((unsigned)nb), dim3(kSNT), 0, s, a);
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(kpp_first_pick_krn<T>, dim3(1), dim3(kPNT), 0, s, a);
  KMD_LAUNCH_CHECK();
  launch_kpp_step<T, 1>(a, 0, s);
  // candidate registers sized to the trial count (2 + floor(ln k) is 2..8
  // for k < 1097)
  auto step = [&](int c) {
    switch (trials) {
      case 1: launch_kpp_step<T, 1>(a, c, s); break;
      case 2: launch_kpp_step<T, 2>(a, c, s); break;
      case 3: launch_kpp_step<T, 3>(a, c, s); break;
      case 4: launch_kpp_step<T, 4>(a, c, s); break;
      case 5: launch_kpp_step<T, 5>(a, c, s); break;
      case 6: launch_kpp_step<T, 6>(a, c, s); break;
      case 7: launch_kpp_step<T, 7>(a, c, s); break;
      case 8: launch_kpp_step<T, 8>(a, c, s); break;
      default: launch_kpp_step<T, 16>(a, c, s); break;
    }
  };
  for (int64_t c = 1; c < k; ++c) step((int)c);
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
