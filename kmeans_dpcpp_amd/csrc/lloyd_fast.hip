// FP32 MFMA fast path of the fused Lloyd step and of the assignment
// (float data, d <= 64, centroids + sums fitting in LDS).
//
// Cross term.  dot[c, s] = sum_f C[f, c] X[f, s] is a (k x d) . (d x n)
// product computed with v_mfma_f32_32x32x2f32: the 32 rows of an MFMA tile
// are centroids, its 32 columns are samples, its K = 2 steps walk the
// features.  In that layout each lane's 16 accumulators belong to a single
// sample (column lane % 32) and 16 centroids, so the argmin is 16 register
// compares plus one exchange with lane ^ 32.
//
// Samples.  A wave processes super-blocks of 32 * S samples: lane l holds
// X[f][s0 + S (l % 32) + q] for its feature parity f = 2 ks + l / 32 and
// q < S, read once from HBM with one S-wide load per feature (the column a
// sample occupies is a free choice of the product, so neighbouring samples
// sit in one lane).  The centroid matrix (zero-padded to 32
// columns and an even number of rows) and the negated half-norms (-inf
// padded) stay in LDS for the life of the persistent block; each A operand
// read from LDS feeds S MFMAs.
//
// Update.  w_s x_s is accumulated into per-block sums in LDS as 64-bit
// fixed-point integers (ds_add_u64): q = rint(w x 2^E_f), with one exponent
// per feature chosen by a pre-pass so that every element fits an int32
// (max|w| max|x_f| 2^E_f < 2^30, a 2^-31 relative resolution per element)
// and the total fits an int64 (n max|w| max|x_f| 2^E_f < 2^61).
// Integer addition is exact and associative, so the sums do not depend on
// the order samples arrive in (no atomics-order nondeterminism) and the only
// rounding is one per element and one at the final conversion.  LDS integer
// atomics run at full LDS rate, unlike ds_add_f32.  If X or w holds a
// non-finite value the pre-pass raises a flag and the kernel accumulates
// plain FP32 with ds_add_f32 instead (so NaN/inf propagate as in the
// reference).  Per-block sums are flushed once: as raw partials into the
// block's own slot (driver: exact, deterministic reduction) or converted to
// FP32 and added with global atomics into private copy (block % n_copies)
// (the fused_lloyd_single_step entry point).
//
// Arithmetic of the labels: the matrix core accumulates -hn[c] + sum_f
// C[f, c] X[f, s] in FP32 (the chain starts from -hn as its C operand), and
// its exact negation is the pseudo-inertia hn[c] - dot; the argmin rule
// (strict <, lowest index on ties) is that of
// src/device_functions.hpp:139-162.  Relative to the sequential FMA chain of
// the reference the scores differ by a few ulps, so labels can differ only
// on sub-ulp margins.  Every FP32 path (tiled, batched, re-score) uses this
// same sequence, so their labels are identical.
#include "fixed_point.hpp"
#include "fp32_tile.hpp"
#include "lloyd_fast.hpp"

namespace kmd {

namespace {

typedef unsigned long long u64;
typedef float f32x4 __attribute__((ext_vector_type(4)));

// 32-sample MFMA column blocks per wave; each A operand read from LDS feeds
// S MFMAs.  One block per wave keeps a wave at ~110 VGPRs, so 16 waves (4
// per SIMD) share a block's LDS copies and hide each other's latencies;
// measured faster than S = 2 at 12 waves in both the MFMA-bound (k = 256)
// and the HBM-bound (k = 8) regime.
constexpr int kS = 1;
constexpr int kMaxD = 64;
constexpr size_t kLdsPerCu = 160 * 1024;

inline int pad32(int64_t k) { return (int)ceil_div<int64_t>(k, 32) * 32; }

// xr[ks][q] = X[2 ks + half][s_first + q] (0 outside the matrix).  FULL:
// d == 2 KS, so every feature row exists.  vec: rows are S-element aligned
// and s_first is a multiple of S, so one S-wide load reads the S samples.
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
        if constexpr (S == 2) {
          typedef float f32x2 __attribute__((ext_vector_type(2)));
          const f32x2 v =
              __builtin_nontemporal_load(reinterpret_cast<const f32x2 *>(row));
          xr[ks][0] = v.x;
          xr[ks][1] = v.y;
        } else {
#pragma unroll
          for (int q = 0; q < S; ++q) xr[ks][q] = __builtin_nontemporal_load(row + q);
        }
      } else {
#pragma unroll
        for (int q = 0; q < S; ++q)
          xr[ks][q] = (s_first + q < n) ? __builtin_nontemporal_load(row + q) : 0.f;
      }
    } else {
#pragma unroll
      for (int q = 0; q < S; ++q) xr[ks][q] = 0.f;
    }
  }
}

// Exact binade-window sums (the reference-API fused step, which runs no
// pass over X for scales).  A finite FP32 value v = m 2^(e - 150) (m the
// 24-bit significand, e the biased exponent) whose binade e lies in
// [lo, lo + W] is the integer m << (e - lo) in units of 2^(lo - 150), with
// no rounding: those values go into the block's int64 LDS sums exactly
// (|m << W| < 2^(24 + W) leaves 2^(39 - W) additions of headroom).  Values
// outside the window (zeros aside), infinities and NaNs are added straight
// into the caller's copy with an FP32 atomic, as the reference adds every
// value.  lo is set per feature so that the window ends kWinMargin binades
// above the largest centroid coordinate of that feature: samples lie near
// their centroids, so nearly every value falls inside.
constexpr int kWinMargin = 4;
constexpr int kWinTopW = 20;   // weights: window up to 2^20

__device__ inline void window_add(u64 *sum, float *fallback, int64_t off,
                                  float v, int lo, int W) {
  const unsigned b = __float_as_uint(v);
  const int sh = (int)((b >> 23) & 0xffu) - lo;
  if ((unsigned)sh <= (unsigned)W) {
    const u64 m = (u64)((b & 0x7fffffu) | 0x800000u) << sh;
    atomicAdd(sum + off, (b >> 31) ? (u64)0 - m : m);
  } else if (v != 0.f) {
    atomic_add(fallback + off, v);
  }
}


// Launch parameters shared by the kernel variants.
struct FastArgs {
  int64_t n;
  int d, k, kpad, vec;
  const float *X, *w, *C, *hn;
  const float *scale;       // [d + 1] fixed-point scales (2^E_f, 2^E_w)
  const int *nonfinite;     // non-zero: FP32 sums (X or w non-finite, or a
                            // range the scales cannot hold)
  int window;               // > 0: no scales; exact binade-window sums with
                            // this many binades (the reference-API step)
  void *labels;
  void *copies;             // u64 [blocks][d*k + k] or float [n_copies][d*k]
  float *sizes_copies;      // float [n_copies][k] (atomic flush only)
  int64_t n_copies;
  int64_t slot0;            // first partial slot of this launch (store flush)
  const int *queue;         // gather mode: [segment][seg_cap] sample indices
  const int *queue_counts;  // gather mode: [segment] entries used
  int n_segments;           // gather mode: <= kMaxSegments
  int64_t seg_cap;
  int *n_rescored;          // gather mode: total queued samples (block 0)
  int n_priv;               // copies of the LDS int64 sums (power of 2; 0 = 1)
  const int *skip;          // when non-null and *skip != 0 the launch does nothing
};

constexpr int kMaxSegments = 1024;

// GATHER: process the samples listed in the segments of a.queue (segment g
// holds a.queue_counts[g] entries from g * seg_cap) instead of 0..n-1: the
// re-scoring of samples the BF16 path left undecided.  Item p of the
// concatenated queue lives in the segment g with off[g] <= p < off[g + 1].
template <int NT, int KS, int S, bool FULL, typename I, bool UPDATE,
          bool STORE_FLUSH, bool GATHER = false>
__global__ void __launch_bounds__(NT) mfma_lloyd_krn(FastArgs a) {
  // an iteration enqueued ahead of a relocation: nothing to compute
  if (a.skip != nullptr && *a.skip != 0) return;
  constexpr int kWavesPerBlock = NT / kWave;
  // The best tile's accumulators are kept in registers (tile_track_best)
  // where the sample operands leave room for 16 more without spilling (not
  // beside the binade-window sums of the atomic-flush form).
  constexpr bool kKeepTile =
      KS * S <= 16 && (FULL || KS * S <= 12) && !(UPDATE && !STORE_FLUSH);
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int64_t n = a.n;
  const int d = a.d, k = a.k, kpad = a.kpad;
  // Cl[((ct KS + ks) 2 + half) 32 + col] = C[2 ks + half][32 ct + col]
  float *Cl = reinterpret_cast<float *>(lds_raw);  // 2 KS kpad floats
  float *Hl = Cl + 2 * KS * kpad;                   // [kpad]
  float *Sc = Hl + kpad;                                  // [d + 1] scales
  // [R][d*k + k] int64 sums: lane l adds into copy l % R, so lanes of one
  // instruction whose samples share a label mostly hit different addresses
  // (few centroids: many lanes per label); the copies are summed at the end
  u64 *S64 = reinterpret_cast<u64 *>(Sc + (kMaxD + 2));
  float *S32 = reinterpret_cast<float *>(S64);            // non-finite path
  const int tid = threadIdx.x;
  const int64_t dk = (int64_t)d * k;
  const int R = (UPDATE && a.n_priv > 1) ? a.n_priv : 1;
  // The atomic-flush form (the reference-API step) always sums in binade
  // windows; the store-flush form (the driver) in fixed point or FP32.
  constexpr bool window = UPDATE && !STORE_FLUSH;
  const bool fixed = UPDATE && !window && a.nonfinite[0] == 0;
  int *Ew = reinterpret_cast<int *>(Sc);   // window mode: lowest binade per row
  int *seg_off = reinterpret_cast<int *>(S64 + R * (dk + k));  // gather: [nseg+1]
  if constexpr (GATHER) {
    // exclusive prefix sum of the segment counts (one wave, 64 per step)
    if (tid < kWave) {
      int carry = 0;
      for (int g0 = 0; g0 < a.n_segments; g0 += kWave) {
        const int g = g0 + tid;
        const int c = (g < a.n_segments) ? a.queue_counts[g] : 0;
        int incl = c;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
          const int t = __shfl_up(incl, o, kWave);
          if (tid >= o) incl += t;
        }
        if (g < a.n_segments) seg_off[g] = carry + incl - c;
        carry += __shfl(incl, kWave - 1, kWave);
      }
      if (tid == 0) {
        seg_off[a.n_segments] = carry;
        if (blockIdx.x == 0) a.n_rescored[0] = carry;
      }
    }
  }

  for (int i = tid; i < 2 * KS * kpad; i += NT) {
    const int cl = i & 31, hf = (i >> 5) & 1, blk = i >> 6;
    const int ks = blk % KS, t = blk / KS;
    const int f = 2 * ks + hf, c = 32 * t + cl;
    Cl[i] = (f < d && c < k) ? a.C[(int64_t)f * k + c] : 0.f;
  }
  for (int c = tid; c < kpad; c += NT)
    Hl[c] = (c < k) ? -a.hn[c] : -__builtin_huge_valf();   // -hn, -inf padded
  if (UPDATE) {
    if (fixed)
      for (int f = tid; f <= d; f += NT) Sc[f] = a.scale[f];
    if constexpr (window)
      for (int f = tid; f <= d; f += NT) Ew[f] = 0;
    for (int64_t i = tid; i < R * (dk + k); i += NT) S64[i] = 0ull;
  }
  __syncthreads();
  if constexpr (window) {
    // the window of feature f ends kWinMargin binades above the largest
    // centroid coordinate of f (that of the weights at 2^kWinTopW)
    for (int64_t i = tid; i < dk; i += NT) {
      const int e = (int)((__float_as_uint(a.C[i]) >> 23) & 0xffu);
      if (e != 0xff) atomicMax(&Ew[i / k], e);
    }
    __syncthreads();
    for (int f = tid; f <= d; f += NT) {
      const int top = f < d ? Ew[f] + kWinMargin : 127 + kWinTopW;
      Ew[f] = min(max(top, a.window + 1), 254) - a.window;
    }
    __syncthreads();
  }

  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;
  const int col = lane & 31;
  const int half = lane >> 5;
  const int n_ct = kpad / 32;
  constexpr int kSuper = 32 * S;
  const int64_t n_items = GATHER ? (int64_t)seg_off[a.n_segments] : n;
  const int64_t n_super = ceil_div<int64_t>(n_items, kSuper);
  const float *X = a.X;
  const float *w = a.w;
  const bool vec = a.vec != 0;
  I *labels = reinterpret_cast<I *>(a.labels);

  // Streaming mode: the next super-block's samples and weights are loaded
  // while this one is scored, so no wave waits on HBM at its start.
  const int64_t sb_stride = (int64_t)gridDim.x * kWavesPerBlock;
  float xn[KS][S], wn[S];
  auto prefetch = [&](int64_t b) {
    const int64_t pf = b * kSuper + (int64_t)S * col;
    load_x<KS, S, FULL>(X, n, d, pf, half, vec, xn);
#pragma unroll
    for (int q = 0; q < S; ++q)
      wn[q] = (UPDATE && pf + q < n) ? __builtin_nontemporal_load(w + pf + q) : 0.f;
  };
  if (!GATHER && (int64_t)blockIdx.x * kWavesPerBlock + wave < n_super)
    prefetch((int64_t)blockIdx.x * kWavesPerBlock + wave);
  for (int64_t sb = (int64_t)blockIdx.x * kWavesPerBlock + wave; sb < n_super;
       sb += sb_stride) {
    const int64_t pos_first = sb * kSuper + (int64_t)S * col;
    int64_t sidx[S];
    bool valid[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
      valid[q] = pos_first + q < n_items;
      sidx[q] = pos_first + q;
      if (GATHER && valid[q]) {
        // last segment whose offset is <= the item's position
        const int p = (int)(pos_first + q);
        int lo = 0, hi = a.n_segments - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (seg_off[mid] <= p)
            lo = mid;
          else
            hi = mid - 1;
        }
        sidx[q] = a.queue[(int64_t)lo * a.seg_cap + (p - seg_off[lo])];
      }
    }
    float xr[KS][S], wv[S];
    if (GATHER) {
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int f = 2 * ks + half;
#pragma unroll
        for (int q = 0; q < S; ++q)
          xr[ks][q] = ((FULL || f < d) && valid[q]) ? X[(int64_t)f * n + sidx[q]]
                                                     : 0.f;
      }
#pragma unroll
      for (int q = 0; q < S; ++q) wv[q] = (UPDATE && valid[q]) ? w[sidx[q]] : 0.f;
    } else {
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int q = 0; q < S; ++q) xr[ks][q] = xn[ks][q];
#pragma unroll
      for (int q = 0; q < S; ++q) wv[q] = wn[q];
      if (sb + sb_stride < n_super) prefetch(sb + sb_stride);
    }

    float bestv[S];
    int bestt[S], besti[S];
    const float *Cw = Cl + lane;
    if (kKeepTile && n_ct == 1) {
      // one tile: its maximum and row straight from the accumulators
      f32x16 acc[S];
      tile_mfma<KS, S>(Cw, Hl + 4 * half, xr, acc);
#pragma unroll
      for (int q = 0; q < S; ++q) {
        const float m = tile_max(acc[q]);
        bestv[q] = -m;
        besti[q] = (bestv[q] < __builtin_huge_valf())
                       ? best_index(0, best_row(acc[q], m), half)
                       : 0;
      }
    } else if constexpr (kKeepTile) {
      float bestm[S];
      f32x16 besta[S];
#pragma unroll
      for (int q = 0; q < S; ++q) {
        bestm[q] = -__builtin_huge_valf();
        bestt[q] = 0;
        besta[q] = f32x16{};
      }
      for (int ct = 0; ct < n_ct; ++ct) {
        f32x16 acc[S];
        tile_mfma<KS, S>(Cw + ct * (KS * 64), Hl + ct * 32 + 4 * half, xr, acc);
        tile_track_best<S>(ct, acc, bestm, bestt, besta);
      }
      // no row of any tile below +inf: centroid 0, as the scan would give
#pragma unroll
      for (int q = 0; q < S; ++q) {
        bestv[q] = -bestm[q];
        besti[q] = (bestv[q] < __builtin_huge_valf())
                       ? best_index(bestt[q], best_row(besta[q], bestm[q]), half)
                       : 0;
      }
    } else {
      int bestr[S];
#pragma unroll
      for (int q = 0; q < S; ++q) {
        bestv[q] = __builtin_huge_valf();
        bestt[q] = 0;
        bestr[q] = 0;
      }
      for (int ct = 0; ct < n_ct; ++ct) {
        f32x16 acc[S];
        tile_mfma<KS, S>(Cw + ct * (KS * 64), Hl + ct * 32 + 4 * half, xr, acc);
        tile_argmin<S>(ct, acc, bestv, bestt, bestr);
      }
#pragma unroll
      for (int q = 0; q < S; ++q)
        besti[q] = (bestv[q] < __builtin_huge_valf())
                       ? best_index(bestt[q], bestr[q], half)
                       : 0;
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
        if (valid[q]) __builtin_nontemporal_store((I)besti[q], labels + sidx[q]);
    }
    if (UPDATE) {
#pragma unroll
      for (int q = 0; q < S; ++q) {
        if (valid[q]) {
          const float ws = wv[q];
          const int lab = besti[q];
          if (fixed) {
            // |w x| 2^E_f < 2^30: the rounded element fits an int32; the
            // sign-extended value is added into the 64-bit sum
            u64 *Sp = S64 + (int64_t)(lane & (R - 1)) * (dk + k);
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
              const int f = 2 * ks + half;
              if (FULL || f < d) {
                const int qv = __float2int_rn((xr[ks][q] * ws) * Sc[f]);
                atomicAdd(&Sp[(int64_t)f * k + lab], (u64)(long long)qv);
              }
            }
            if (half == 0)
              atomicAdd(&Sp[dk + lab],
                        (u64)(long long)__float2int_rn(ws * Sc[d]));
          } else if constexpr (window) {
            u64 *Sp = S64 + (int64_t)(lane & (R - 1)) * (dk + k);
            const int64_t copy = blockIdx.x % a.n_copies;
            float *cg = reinterpret_cast<float *>(a.copies) + copy * dk;
#pragma unroll 4
            for (int ks = 0; ks < KS; ++ks) {
              const int f = 2 * ks + half;
              if (FULL || f < d)
                window_add(Sp, cg, (int64_t)f * k + lab, xr[ks][q] * ws, Ew[f],
                           a.window);
            }
            if (half == 0)
              window_add(Sp + dk, a.sizes_copies + copy * k, lab, ws, Ew[d],
                         a.window);
          } else {
            // FP32 sums, spread over the R copies like the integer ones
            float *Sp = S32 + (int64_t)(lane & (R - 1)) * (dk + k);
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
              const int f = 2 * ks + half;
              if (FULL || f < d)
                atomic_add(&Sp[(int64_t)f * k + lab], xr[ks][q] * ws);
            }
            if (half == 0) atomic_add(&Sp[dk + lab], ws);
          }
        }
      }
    }
  }
  if (UPDATE) {
    __syncthreads();
    if (STORE_FLUSH) {
      // raw partials (int64 fixed point, or FP32 on the non-finite path)
      // into this block's slot
      if (fixed) {
        u64 *cp = reinterpret_cast<u64 *>(a.copies) +
                  (a.slot0 + blockIdx.x) * (dk + k);
        for (int64_t i = tid; i < dk + k; i += NT) {
          u64 t = S64[i];
          for (int r = 1; r < R; ++r) t += S64[r * (dk + k) + i];
          cp[i] = t;
        }
      } else {
        float *cp = reinterpret_cast<float *>(a.copies) +
                    (a.slot0 + blockIdx.x) * (dk + k);
        for (int64_t i = tid; i < dk + k; i += NT) {
          float t = S32[i];
          for (int r = 1; r < R; ++r) t += S32[r * (dk + k) + i];
          cp[i] = t;
        }
      }
    } else {
      const int64_t copy = blockIdx.x % a.n_copies;
      float *cp = reinterpret_cast<float *>(a.copies) + copy * dk;
      float *sp = a.sizes_copies + copy * k;
      for (int64_t i = tid; i < dk + k; i += NT) {
        float v;
        if (fixed) {
          const int f = (int)(i / k);  // f == d for the counts
          u64 t = S64[i];
          for (int r = 1; r < R; ++r) t += S64[r * (dk + k) + i];
          v = (float)((double)(long long)t / (double)Sc[f]);
        } else if constexpr (window) {
          const int f = (int)(i / k);  // f == d for the counts
          u64 t = S64[i];
          for (int r = 1; r < R; ++r) t += S64[r * (dk + k) + i];
          v = (float)ldexp((double)(long long)t, Ew[f] - 150);
        } else {
          v = S32[i];
          for (int r = 1; r < R; ++r) v += S32[r * (dk + k) + i];
        }
        if (v != 0.f) {
          if (i < dk)
            atomic_add(&cp[i], v);
          else
            atomic_add(&sp[i - dk], v);
        }
      }
    }
  }
}

// ---- few centroids (k <= 32) -----------------------------------------------
//
// With k <= 16 the 32-row tile above computes 2-8 x the products it needs,
// up to k = 32 it needs an exchange between lanes for the argmin,
// and its per-block sums take most of a wave's 64 atomic adds into a handful
// of addresses.  This kernel gives each lane one sample and all of its
// centroids instead:
//
// Cross term.  v_mfma_f32_4x4x1f32 runs 16 independent 4 x 4 products with
// K = 1: block b = lane / 4 multiplies the 4-vector C[f][4 g .. 4 g + 3]
// (the same in every block) by X[f][s0 + 4 b .. 4 b + 3], so lane l ends
// with the scores of sample s0 + l against centroids 4 g .. 4 g + 3 in its
// 4 accumulators, for each of G = ceil(k / 4) groups.  A chain of FP32 MFMAs
// (any shape) rounds exactly like fmaf(c, x, acc) applied feature after
// feature in ascending order, so the chain started from -hn gives the same
// bits as the 32x32x2 tile chain of every other FP32 path: same scores,
// same labels.  The argmin over the lane's 4 G scores needs no exchange
// between lanes.
//
// Update.  The int64 fixed-point sums live in LDS as [label][feature][copy]
// with the copy index (lane % R) fastest: in one ds_add_u64 the 64 lanes
// add into words whose offsets differ by their copy index, whatever their
// labels.  With R = 32 (k <= 8 at d <= 32) no two lanes of an instruction
// share an address except lanes l and l + 32, and the atomics run at the
// LDS's full width instead of serialising on the few labels; larger k or d
// leave room for fewer copies (R = 4-16), each shared by 64 / R lanes.
constexpr int kSmallK = 32;

// Per feature-row bucket DB and centroid groups G: threads per block,
// copies of the sums (R k (DB + 1) int64 beside the centroids and the label
// stage must fit the LDS; the offsets f R of a lane's adds are then
// compile-time immediates) and super-blocks of labels a wave stages in LDS.
template <int DB, int G>
constexpr int smallk_threads() {
  return DB <= 32 ? ((DB == 32 && G == 8) ? 768 : 1024) : 512;
}
template <int DB, int G>
constexpr int smallk_copies() {
  return DB <= 32 ? (G == 8 ? 8 : G == 4 ? 16 : 32) : (G == 8 ? 4 : G == 4 ? 8 : 16);
}
// Labels are held in LDS, one byte per sample (k <= 32), and written out
// when kLabelStage super-blocks have been scored and at the end: HBM writes
// interleaved with the stream of reads cost far more than their bytes (at
// n = 10 M, d = 32 the label stores of every super-block added 15-40 us to
// a 0.25 ms step), so a wave writes its labels in one burst after its last
// load wherever n allows.
constexpr int kLabelStage = 64;
// cache-policy operand of the sample loads: non-temporal (read once per step)
constexpr int kLoadNT = 2;
// Word 3 of a raw buffer descriptor on gfx9-family devices (untyped dword
// access, no swizzle).
constexpr int kRsrcWord3 = 0x00020000;

// acc += C[f][4 g + r] x (row r of every 4x4 block): CBSZ = 4 makes the 16
// blocks one broadcast group and ABID = g hands them block g's A values,
// i.e. lanes 4 g .. 4 g + 3, which hold centroids 4 g .. 4 g + 3.
template <int Gi>
__device__ inline f32x4 mfma4_group(float a, float x, f32x4 acc) {
  return __builtin_amdgcn_mfma_f32_4x4x1f32(a, x, acc, 4, Gi, 0);
}
// one feature's products for groups Gi .. G - 1
template <int G, int Gi = 0>
__device__ inline void mfma4_groups(float a, float x, f32x4 (&acc)[G]) {
  acc[Gi] = mfma4_group<Gi>(a, x, acc[Gi]);
  if constexpr (Gi + 1 < G) mfma4_groups<G, Gi + 1>(a, x, acc);
}

// window_add for the few-centroid layout: the LDS word and the caller's
// FP32 element of an out-of-window value are addressed separately.  `real`
// is false for a padded feature row (x = 0, so v is 0, or NaN under an
// infinite weight: never inside a window), whose value has no element in
// the caller's copy and is dropped.
__device__ inline void window_add_at(u64 *word, float *fb, int64_t fb_off,
                                     float v, int lo, int W, bool real) {
  const unsigned b = __float_as_uint(v);
  const int sh = (int)((b >> 23) & 0xffu) - lo;
  if ((unsigned)sh <= (unsigned)W) {
    const u64 m = (u64)((b & 0x7fffffu) | 0x800000u) << sh;
    atomicAdd(word, (b >> 31) ? (u64)0 - m : m);
  } else if (v != 0.f && real) {
    atomic_add(fb + fb_off, v);
  }
}

// UPDATE && !WINDOW: the driver's step (fixed-point sums, or FP32 for
// non-finite inputs; raw partials into the block's slot).  WINDOW: the
// reference's fused step (exact binade-window sums, no pre-pass; totals and
// out-of-window values added into the caller's private copies).
template <int DB, int G, typename I, bool UPDATE, bool WINDOW = false>
__global__ void __launch_bounds__((smallk_threads<DB, G>()))
    smallk_lloyd_krn(FastArgs a) {
  static_assert(!WINDOW || UPDATE, "the window sums belong to the update");
  if (a.skip != nullptr && *a.skip != 0) return;
  constexpr int NT = smallk_threads<DB, G>();
  constexpr int R = smallk_copies<DB, G>();
  constexpr int kWavesPerBlock = NT / kWave;
  constexpr int KP = 4 * G;
  // scales and -hn in scalar registers where they fit
  constexpr bool kScalarConsts = DB <= 32;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int64_t n = a.n;
  const int d = a.d, k = a.k;
  // Ca[c DB + f] = C[f][c] (0 outside the matrix): lane c < KP reads its
  // centroid's A operands of four consecutive features with one 16-byte read
  float *Ca = reinterpret_cast<float *>(lds_raw);
  float *Hl = Ca + KP * DB;                          // [KP] -hn, -inf padded
  float *Sc = Hl + KP;                               // [DB + 4] scales
  u64 *S64 = reinterpret_cast<u64 *>(Sc + DB + 4);   // [k][DB + 1][R]
  float *S32 = reinterpret_cast<float *>(S64);       // non-finite path
  const int tid = threadIdx.x;
  const int n_sums = UPDATE ? k * (DB + 1) * R : 0;
  // [wave][kLabelStage][64] labels of the wave's staged super-blocks
  unsigned char *stage_all = reinterpret_cast<unsigned char *>(S64 + n_sums);
  const bool fixed = UPDATE && !WINDOW && a.nonfinite[0] == 0;
  int *Ew = reinterpret_cast<int *>(Sc);   // WINDOW: lowest binade per row
  const int W = a.window;

  const int lane = tid & (kWave - 1);
  // uniform in the wave (the compiler cannot see that of tid / 64), so the
  // per-row buffer descriptors below stay in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
  const int64_t n_super = ceil_div<int64_t>(n, kWave);
  const int64_t sb_stride = (int64_t)gridDim.x * kWavesPerBlock;

  // Samples of super-block b: row f is read through a buffer descriptor
  // that covers the block's valid samples of that row and nothing else, so
  // lanes past the last sample and rows f >= d read 0 (out-of-range buffer
  // loads return zero) with no branch: the padded rows then add 0 x 0 to
  // the chains (their C entries are 0 too) and 0 to sum slots that are
  // never flushed.
  auto load = [&](int64_t b, float (&x)[DB], float &wv) {
    // d and n re-enter through an empty asm each time, so that the compiler
    // walks the rows with one running address instead of hoisting DB row
    // bases and row tests out of the loop (scalar registers run out)
    int dd = d;
    int64_t nn = n;
    asm volatile("" : "+s"(dd), "+s"(nn));
    const int64_t s0 = b * kWave;
    const int bytes = (int)(min<int64_t>(kWave, nn - s0) * 4);
    const float *p = a.X + s0;
#pragma unroll
    for (int f = 0; f < DB; ++f) {
      const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<float *>(p), (short)0, f < dd ? bytes : 0, kRsrcWord3);
      x[f] = __uint_as_float(
          __builtin_amdgcn_raw_buffer_load_b32(r, lane * 4, 0, kLoadNT));
      if (f + 1 < dd) p += nn;
    }
    if (UPDATE) {
      const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<float *>(a.w + s0), (short)0, bytes, kRsrcWord3);
      wv = __uint_as_float(
          __builtin_amdgcn_raw_buffer_load_b32(r, lane * 4, 0, kLoadNT));
    }
  };

  // two register buffers: the next super-block is in flight while this one
  // is scored (a third would need more loads in flight than the memory
  // counter can track); the first one's loads are issued before the block
  // sets up its LDS
  float xa[DB], xb[DB], wa = 0.f, wb = 0.f;
  int64_t sb = (int64_t)blockIdx.x * kWavesPerBlock + wave;
  if (sb < n_super) load(sb, xa, wa);

  for (int i = tid; i < KP * DB; i += NT) {
    const int c = i / DB, f = i % DB;
    Ca[i] = (f < d && c < k) ? a.C[(int64_t)f * k + c] : 0.f;
  }
  for (int c = tid; c < KP; c += NT)
    Hl[c] = (c < k) ? -a.hn[c] : -__builtin_huge_valf();
  if (UPDATE) {
    // Sc[f] = the scale of feature f < d, 0 for padded rows; Sc[DB] = the
    // weights' scale
    for (int f = tid; f < DB + 4; f += NT) {
      if constexpr (WINDOW)
        Ew[f] = 0;
      else
        Sc[f] = !fixed ? 0.f : f < d ? a.scale[f] : f == DB ? a.scale[d] : 0.f;
    }
    for (int i = tid; i < n_sums; i += NT) S64[i] = 0ull;
  }
  __syncthreads();
  if constexpr (WINDOW) {
    // the window of feature f ends kWinMargin binades above the largest
    // centroid coordinate of f, that of the weights (slot DB) at 2^kWinTopW
    const int64_t dk = (int64_t)d * k;
    for (int64_t i = tid; i < dk; i += NT) {
      const int e = (int)((__float_as_uint(a.C[i]) >> 23) & 0xffu);
      if (e != 0xff) atomicMax(&Ew[i / k], e);
    }
    __syncthreads();
    for (int f = tid; f <= DB; f += NT) {
      const int top = f < d ? Ew[f] + kWinMargin : 127 + kWinTopW;
      Ew[f] = min(max(top, W + 1), 254) - W;
    }
    __syncthreads();
  }

  I *labels = reinterpret_cast<I *>(a.labels);
  unsigned char *stage = stage_all + wave * (kLabelStage * kWave);
  auto uniform = [](float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
  };
  float hs[kScalarConsts ? KP : 1], ss[kScalarConsts ? DB + 1 : 1];
  if constexpr (kScalarConsts) {
#pragma unroll
    for (int c = 0; c < KP; ++c) hs[c] = uniform(Hl[c]);
    if (UPDATE) {
#pragma unroll
      for (int f = 0; f <= DB; ++f) ss[f] = uniform(Sc[f]);   // WINDOW: Ew bits
    }
  }
  const int64_t copy = UPDATE ? blockIdx.x % a.n_copies : 0;


  // slot: the super-block's place in the stage
  auto body = [&](int64_t sb, int slot, const float (&xr)[DB], float ws) {
    const int64_t s = sb * kWave + lane;
    const bool valid = s < n;
    f32x4 acc[G];   // the chains start from -hn
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if constexpr (kScalarConsts) {
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[g][r] = hs[4 * g + r];
      } else {
        acc[g] = *reinterpret_cast<const f32x4 *>(Hl + 4 * g);
      }
    }
#pragma unroll
    for (int f4 = 0; f4 < DB; f4 += 4) {
      // only lanes c < KP supply A operands (the broadcast reads no other)
      f32x4 av = {};
      if (lane < KP) av = *reinterpret_cast<const f32x4 *>(Ca + lane * DB + f4);
#pragma unroll
      for (int e = 0; e < 4; ++e) mfma4_groups<G>(av[e], xr[f4 + e], acc);
    }
    // largest acc (smallest score hn - c.x) over the KP rows, NaNs passed
    // over, -inf when there is none; then its lowest centroid
    float m = -__builtin_huge_valf();
#pragma unroll
    for (int g = 0; g < G; ++g) {
      m = vmax3(m, acc[g][0], acc[g][1]);
      m = vmax3(m, acc[g][2], acc[g][3]);
    }
    int lab = 0;
#pragma unroll
    for (int c = KP - 1; c >= 0; --c) lab = (acc[c / 4][c % 4] == m) ? c : lab;
    lab = (-m < __builtin_huge_valf()) ? lab : 0;
    stage[slot * kWave + lane] = (unsigned char)lab;

    if (UPDATE && valid) {
      const int base = lab * (DB + 1) * R + (lane & (R - 1));
      if constexpr (WINDOW) {
        u64 *Sp = S64 + base;
        float *cg = reinterpret_cast<float *>(a.copies) + copy * (int64_t)d * k;
#pragma unroll
        for (int f = 0; f < DB; ++f) {
          const int lo = kScalarConsts ? __float_as_int(ss[f]) : Ew[f];
          window_add_at(Sp + f * R, cg, (int64_t)f * k + lab, xr[f] * ws, lo, W,
                        f < d);
        }
        const int low = kScalarConsts ? __float_as_int(ss[DB]) : Ew[DB];
        window_add_at(Sp + DB * R, a.sizes_copies + copy * k, lab, ws, low, W, true);
      } else if (fixed) {
        u64 *Sp = S64 + base;
#pragma unroll
        for (int f4 = 0; f4 < DB; f4 += 4) {
          f32x4 sc;
          if constexpr (kScalarConsts) {
#pragma unroll
            for (int e = 0; e < 4; ++e) sc[e] = ss[f4 + e];
          } else {
            sc = *reinterpret_cast<const f32x4 *>(Sc + f4);
          }
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int qv = __float2int_rn((xr[f4 + e] * ws) * sc[e]);
            atomicAdd(Sp + (f4 + e) * R, (u64)(long long)qv);
          }
        }
        const float scw = kScalarConsts ? ss[DB] : Sc[DB];
        atomicAdd(Sp + DB * R, (u64)(long long)__float2int_rn(ws * scw));
      } else {
        float *Sp = S32 + base;
#pragma unroll
        for (int f = 0; f < DB; ++f) atomic_add(Sp + f * R, xr[f] * ws);
        atomic_add(Sp + DB * R, ws);
      }
    }
  };

  // the labels of super-blocks first, first + stride, ... (cnt of them),
  // each lane its own column of the stage (no exchange between lanes)
  auto flush = [&](int64_t first, int cnt) {
    for (int j = 0; j < cnt; ++j) {
      const int64_t s = (first + j * sb_stride) * kWave + lane;
      if (s < n) __builtin_nontemporal_store((I)stage[j * kWave + lane], labels + s);
    }
  };
  int64_t staged_first = sb;
  int slot = 0;
  auto next_slot = [&](int64_t done) {
    if (++slot == kLabelStage) {
      flush(staged_first, kLabelStage);
      staged_first = done + sb_stride;
      slot = 0;
    }
  };
  for (; sb < n_super; sb += 2 * sb_stride) {
    if (sb + sb_stride < n_super) load(sb + sb_stride, xb, wb);
    body(sb, slot, xa, wa);
    next_slot(sb);
    if (sb + sb_stride >= n_super) break;
    if (sb + 2 * sb_stride < n_super) load(sb + 2 * sb_stride, xa, wa);
    body(sb + sb_stride, slot, xb, wb);
    next_slot(sb + sb_stride);
  }
  flush(staged_first, slot);

  if (UPDATE) {
    // raw partials into this block's slot, in the [f][label] layout of the
    // other fused-step kernels (f == d: the weight sums); WINDOW: the
    // totals into the caller's private copy
    __syncthreads();
    const int dk1 = (d + 1) * k;
    if constexpr (WINDOW) {
      float *cg = reinterpret_cast<float *>(a.copies) + copy * (int64_t)d * k;
      float *sg = a.sizes_copies + copy * k;
      for (int i = tid; i < dk1; i += NT) {
        const int f = i / k, c = i % k;
        const int fs = f < d ? f : DB;
        const u64 *src = S64 + (c * (DB + 1) + fs) * R;
        u64 t = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) t += src[r];
        const float v = (float)ldexp((double)(long long)t, Ew[fs] - 150);
        if (v != 0.f) {
          if (f < d)
            atomic_add(&cg[(int64_t)f * k + c], v);
          else
            atomic_add(&sg[c], v);
        }
      }
    } else if (fixed) {
      u64 *cp = reinterpret_cast<u64 *>(a.copies) + (a.slot0 + blockIdx.x) * dk1;
      for (int i = tid; i < dk1; i += NT) {
        const int f = i / k, c = i % k;
        const u64 *src = S64 + (c * (DB + 1) + (f < d ? f : DB)) * R;
        u64 t = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) t += src[r];
        cp[i] = t;
      }
    } else {
      float *cp = reinterpret_cast<float *>(a.copies) + (a.slot0 + blockIdx.x) * dk1;
      for (int i = tid; i < dk1; i += NT) {
        const int f = i / k, c = i % k;
        const float *src = S32 + (c * (DB + 1) + (f < d ? f : DB)) * R;
        float t = src[0];
#pragma unroll
        for (int r = 1; r < R; ++r) t += src[r];
        cp[i] = t;
      }
    }
  }
}

// ---- fixed-point scales ------------------------------------------------------

// The range statistics of fixed_point.hpp for row f = blockIdx.y of X (f ==
// d: the weights): maximum bit pattern of |value| (for non-negative floats
// the bit pattern orders like the value; any NaN is above +inf), exponent
// sum and count of the normal values, and for the weights the complement of
// the smallest non-zero |w|.
__global__ void __launch_bounds__(256)
    absmax_bits_krn(int64_t n, int d, const float *__restrict__ X,
                    const float *__restrict__ w, unsigned *maxbits,
                    unsigned long long *estat) {
  const int f = blockIdx.y;  // 0..d (d = weights)
  const float *row = (f < d) ? X + (int64_t)f * n : w;
  const bool weights = f == d;
  unsigned m = 0u, esum = 0u, ecnt = 0u, wmin_inv = 0u;
  auto acc = [&](unsigned v) {
    range_accumulate(v, m, esum, ecnt);
    const unsigned b = v & 0x7fffffffu;
    if (weights && b != 0u) wmin_inv = max(wmin_inv, ~b);
  };
  if ((reinterpret_cast<uintptr_t>(row) & 15u) == 0) {
    // 16-byte loads, four in flight per thread; the tail element by element
    const int64_t n4 = n / 4;
    const uint4 *r4 = reinterpret_cast<const uint4 *>(row);
    const int64_t stride = (int64_t)gridDim.x * 256;
    int64_t i = blockIdx.x * 256ll + threadIdx.x;
    for (; i + 3 * stride < n4; i += 4 * stride) {
      const uint4 a = r4[i], b = r4[i + stride], c = r4[i + 2 * stride],
                  e = r4[i + 3 * stride];
      acc(a.x), acc(a.y), acc(a.z), acc(a.w);
      acc(b.x), acc(b.y), acc(b.z), acc(b.w);
      acc(c.x), acc(c.y), acc(c.z), acc(c.w);
      acc(e.x), acc(e.y), acc(e.z), acc(e.w);
    }
    for (; i < n4; i += stride) {
      const uint4 a = r4[i];
      acc(a.x), acc(a.y), acc(a.z), acc(a.w);
    }
    for (int64_t j = 4 * n4 + blockIdx.x * 256ll + threadIdx.x; j < n; j += stride)
      acc(__float_as_uint(row[j]));
  } else {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
      acc(__float_as_uint(row[i]));
  }
  unsigned long long es = esum, ec = ecnt;
  for (int o = kWave / 2; o > 0; o >>= 1) {
    m = max(m, (unsigned)__shfl_xor((int)m, o, kWave));
    wmin_inv = max(wmin_inv, (unsigned)__shfl_xor((int)wmin_inv, o, kWave));
    es += __shfl_xor(es, o, kWave);
    ec += __shfl_xor(ec, o, kWave);
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    atomicMax(&maxbits[f], m);
    if (weights && wmin_inv != 0u) atomicMax(&maxbits[d + 1], wmin_inv);
    if (ec != 0) {
      atomicAdd(&estat[2 * f], es);
      atomicAdd(&estat[2 * f + 1], ec);
    }
  }
}

// words[0 : count] = 0 (a kernel rather than a memset node, so that callers
// may capture the launches in a graph)
__global__ void zero_words_krn(unsigned *words, int64_t count) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * 256)
    words[i] = 0u;
}

// Scales and flag from the statistics (fixed_point.hpp), one thread.
__global__ void scales_krn(int64_t n, int d, const unsigned *maxbits,
                           const unsigned long long *estat, float *scale,
                           int *flag) {
  fixed_point_scales_from_bits(n, d, maxbits, estat, scale, flag);
}

// ---- launch planning -----------------------------------------------------

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

inline size_t lds_bytes(int64_t d, int64_t k, bool update, int copies = 1) {
  const int kpad = pad32(k);
  const int KS = ks_bucket(d);
  size_t b = sizeof(float) * ((size_t)2 * KS * kpad + kpad);
  if (update)
    b += sizeof(float) * (kMaxD + 2) +
         sizeof(u64) * (size_t)copies * (d + 1) * k;
  return b;
}

// Copies of the LDS sums the fused step keeps (up to 8, as LDS allows: the
// kernel runs one block per CU).
inline int sum_copies(int64_t d, int64_t k) {
  int r = 1;
  while (r < 8 && lds_bytes(d, k, true, 2 * r) <= kLdsPerCu) r *= 2;
  return r;
}

using KernFn = void (*)(FastArgs);

template <int NT, typename I, bool UPDATE, bool STORE, int S = kS>
KernFn pick_nt(int64_t d) {
  const int ks = ks_bucket(d);
  const bool full = (d == 2 * ks);
#define KMD_KS_CASE(KS)                                                   \
  case KS:                                                                \
    return full ? &mfma_lloyd_krn<NT, KS, S, true, I, UPDATE, STORE>      \
                : &mfma_lloyd_krn<NT, KS, S, false, I, UPDATE, STORE>;
  switch (ks) {
    KMD_KS_CASE(2)
    KMD_KS_CASE(4)
    KMD_KS_CASE(8)
    KMD_KS_CASE(12)
    KMD_KS_CASE(16)
    KMD_KS_CASE(24)
    KMD_KS_CASE(32)
    default:
      return nullptr;
  }
#undef KMD_KS_CASE
}

// Threads per block of the MFMA kernels: 16 waves share one copy of the
// centroids and of the per-block sums in LDS.
constexpr int kThreads = 1024;

// Launch shapes of the fixed-point fused step that the autotuner chooses
// from (tuned_shape()): threads per block x 32-sample column blocks per wave.
// Every shape computes the same per-pair arithmetic, so the same labels and
// sums; they differ in how many waves share a block's LDS and how many
// samples each wave keeps in registers.
struct FastShape {
  int threads, s;
};
constexpr FastShape kFastShapes[] = {{1024, 1}, {768, 2}, {512, 2}};
constexpr int kNumFastShapes = 3;

// A resolved launch: kernel variant, block size, dynamic LDS and the
// persistent grid (occupancy-limited, never more blocks than work).
struct MfmaLaunch {
  KernFn fn = nullptr;
  int threads = kThreads;
  int s = kS;      // 32-sample column blocks per wave
  int blocks = 0;
  size_t lds = 0;
  int copies = 1;  // copies of the LDS sums (UPDATE)
  bool smallk = false;   // the few-centroid kernel (64-sample super-blocks)
};

template <typename I, bool UPDATE, bool STORE>
KernFn pick_shape(int64_t d, int shape) {
  if constexpr (UPDATE && STORE) {
    if (shape == 1) return pick_nt<768, I, true, true, 2>(d);
    if (shape == 2) return pick_nt<512, I, true, true, 2>(d);
  }
  return pick_nt<kThreads, I, UPDATE, STORE>(d);
}

// ---- few-centroid launches (smallk_lloyd_krn) ----

// Feature rows a lane holds: d rounded up to 8, 16, 32 or 64.
inline int smallk_db(int64_t d) {
  return d <= 8 ? 8 : d <= 16 ? 16 : d <= 32 ? 32 : 64;
}

// Groups of 4 centroids the kernel computes (its template G): 1, 2 or 4.
inline int smallk_groups(int64_t k) { return k <= 4 ? 1 : k <= 8 ? 2 : k <= 16 ? 4 : 8; }

// The kernel's threads per block (smallk_threads<DB, G>).
inline int smallk_threads_of(int64_t d, int64_t k) {
  const int db = smallk_db(d), g = smallk_groups(k);
  if (db == 32 && g == 8) return smallk_threads<32, 8>();
  return db <= 32 ? smallk_threads<16, 1>() : smallk_threads<64, 1>();
}

inline size_t smallk_lds(int64_t d, int64_t k, bool update, int copies) {
  const int db = smallk_db(d);
  const int kp = 4 * smallk_groups(k);   // the kernel's KP rows
  const int threads = smallk_threads_of(d, k);
  size_t b = sizeof(float) * ((size_t)kp * db + kp + db + 4);
  if (update) b += sizeof(u64) * (size_t)k * (db + 1) * copies;
  return b + (size_t)(threads / kWave) * kLabelStage * kWave;   // label stage
}

// The kernel's copies of the sums (smallk_copies<DB, G>).
inline int smallk_copies_of(int64_t d, int64_t k) {
  const int g = smallk_groups(k);
  if (smallk_db(d) <= 32)
    return g == 8 ? smallk_copies<32, 8>() : g == 4 ? smallk_copies<32, 4>()
                                                    : smallk_copies<32, 2>();
  return g == 8 ? smallk_copies<64, 8>() : g == 4 ? smallk_copies<64, 4>()
                                                  : smallk_copies<64, 2>();
}

// KMD_NO_SMALLK=1 sends k <= 32 to the 32-row tile kernel (the two compute
// the same labels and sums; the switch exists to compare them).
inline bool smallk_enabled() {
  const char *e = std::getenv("KMD_NO_SMALLK");
  return !(e != nullptr && e[0] == '1' && e[1] == '\0');
}

template <typename I, bool UPDATE, bool WINDOW>
KernFn pick_smallk(int64_t d, int64_t k) {
  const int g = smallk_groups(k);
  const int db = smallk_db(d);
#define KMD_SK_CASE(DB)                                                \
  case DB:                                                             \
    return g == 1   ? &smallk_lloyd_krn<DB, 1, I, UPDATE, WINDOW>      \
           : g == 2 ? &smallk_lloyd_krn<DB, 2, I, UPDATE, WINDOW>      \
           : g == 4 ? &smallk_lloyd_krn<DB, 4, I, UPDATE, WINDOW>      \
                    : &smallk_lloyd_krn<DB, 8, I, UPDATE, WINDOW>;
  switch (db) {
    KMD_SK_CASE(8)
    KMD_SK_CASE(16)
    KMD_SK_CASE(32)
    KMD_SK_CASE(64)
    default:
      return nullptr;
  }
#undef KMD_SK_CASE
}

template <typename I, bool UPDATE, bool WINDOW = false>
MfmaLaunch plan_smallk(int64_t n, int64_t d, int64_t k) {
  MfmaLaunch L;
  L.smallk = true;
  // the kernel's lane-keyed copies of the sums (fixed per bucket)
  L.copies = UPDATE ? smallk_copies_of(d, k) : 1;
  L.lds = smallk_lds(d, k, UPDATE, L.copies);
  if (L.lds > kLdsPerCu) return L;
  L.threads = smallk_threads_of(d, k);
  L.s = 1;
  L.fn = pick_smallk<I, UPDATE, WINDOW>(d, k);
  if (L.fn == nullptr) return L;
  allow_full_lds(reinterpret_cast<const void *>(L.fn));
  int per_cu = 0;
  KMD_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void *>(L.fn), L.threads, L.lds));
  per_cu = apply_blocks_cap(per_cu);
  if (per_cu < 1) {
    L.fn = nullptr;
    return L;
  }
  const int64_t need = ceil_div<int64_t>(
      ceil_div<int64_t>(n > 0 ? n : 1, kWave), L.threads / kWave);
  int64_t blocks = (int64_t)device_cu_count() * per_cu;
  if (blocks > need) blocks = need;
  L.blocks = (int)(blocks < 1 ? 1 : blocks);
  return L;
}

template <typename I, bool UPDATE, bool STORE>
MfmaLaunch plan_launch(int64_t n, int64_t d, int64_t k) {
  MfmaLaunch L;
  if (d < 1 || d > kMaxD || k < 1 || k > (1 << 20)) return L;
  // few centroids: the driver's step (store flush), the assignment and the
  // reference-API step (atomic flush, binade windows)
  if (k <= kSmallK && smallk_enabled()) {
    L = plan_smallk<I, UPDATE, UPDATE && !STORE>(n, d, k);
    if (L.fn != nullptr) return L;
    L = MfmaLaunch{};
  }
  L.copies = UPDATE ? sum_copies(d, k) : 1;
  L.lds = lds_bytes(d, k, UPDATE, L.copies);
  if (L.lds > kLdsPerCu) return L;
  int shape = (UPDATE && STORE) ? tuned_shape() : 0;
  if (shape < 0 || shape >= kNumFastShapes) shape = 0;
  L.threads = kFastShapes[shape].threads;
  L.s = kFastShapes[shape].s;
  L.fn = pick_shape<I, UPDATE, STORE>(d, shape);
  if (L.fn == nullptr) return L;
  static thread_local KernFn attr_done[128] = {};
  bool seen = false;
  for (KernFn f : attr_done) seen |= (f == L.fn);
  if (!seen) {
    KMD_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(L.fn),
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)kLdsPerCu));
    for (KernFn &f : attr_done)
      if (f == nullptr) {
        f = L.fn;
        break;
      }
  }
  int per_cu = 0;
  KMD_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void *>(L.fn), L.threads, L.lds));
  per_cu = apply_blocks_cap(per_cu);
  if (per_cu < 1) {
    L.fn = nullptr;
    return L;
  }
  const int64_t waves_per_block = L.threads / kWave;
  const int64_t n_super = ceil_div<int64_t>(n > 0 ? n : 1, 32 * L.s);
  int64_t blocks = (int64_t)device_cu_count() * per_cu;
  const int64_t need = ceil_div<int64_t>(n_super, waves_per_block);
  if (blocks > need) blocks = need;
  if (blocks < 1) blocks = 1;
  L.blocks = (int)blocks;
  return L;
}

inline void run(const MfmaLaunch &L, const FastArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(L.fn, dim3(L.blocks), dim3(L.threads), L.lds, s, a);
  KMD_LAUNCH_CHECK();
}

inline bool vec_ok(const float *X, int64_t n, int64_t S = kS) {
  return (n % S == 0) && ((reinterpret_cast<uintptr_t>(X) & (4 * S - 1)) == 0);
}

inline FastArgs make_args(int64_t n, int64_t d, int64_t k, const float *X,
                          const float *w, const float *C, const float *hn,
                          void *labels) {
  FastArgs a{};
  a.n = n;
  a.d = (int)d;
  a.k = (int)k;
  a.kpad = pad32(k);
  a.vec = vec_ok(X, n) ? 1 : 0;
  a.X = X;
  a.w = w;
  a.C = C;
  a.hn = hn;
  a.labels = labels;
  return a;
}

}  // namespace

size_t fixed_point_stats_bytes(int64_t d) {
  const size_t est_off = (sizeof(unsigned) * range_maxbits_count(d) + 15) & ~size_t(15);
  return est_off + sizeof(unsigned long long) * range_estat_count(d);
}

void launch_fixed_point_scales(int64_t n, int64_t d, const float *X,
                               const float *w, float *scale, int *nonfinite,
                               hipStream_t s, void *stats_scratch) {
  const size_t est_off = (sizeof(unsigned) * range_maxbits_count(d) + 15) & ~size_t(15);
  const size_t bytes = fixed_point_stats_bytes(d);
  TempBuffer own(stats_scratch ? 0 : bytes, s);
  char *stats = stats_scratch ? static_cast<char *>(stats_scratch) : own.as<char>();
  hipLaunchKernelGGL(zero_words_krn, dim3((unsigned)ceil_div<int64_t>(bytes / 4, 256)),
                     dim3(256), 0, s, reinterpret_cast<unsigned *>(stats),
                     (int64_t)(bytes / 4));
  KMD_LAUNCH_CHECK();
  unsigned *maxbits = reinterpret_cast<unsigned *>(stats);
  unsigned long long *estat = reinterpret_cast<unsigned long long *>(stats + est_off);
  // blocks per row: each wave ends with same-address atomics, which
  // serialise (about 0.1-0.2 us apiece per address), so few blocks per row
  // -- 16 at d = 32, enough to fill the chip for small d -- each streaming
  // a long stretch of it: 0.23 ms at n = 10 M, d = 32, against 0.41 ms with
  // 249 blocks per row
  int gx = (int)ceil_div<int64_t>(n > 0 ? n : 1, 256 * 32);
  const int cap = std::max(16, 256 / (int)(d + 1));
  if (gx > cap) gx = cap;
  if (gx < 1) gx = 1;
  if (n > 0) {
    hipLaunchKernelGGL(absmax_bits_krn, dim3(gx, (unsigned)(d + 1)), dim3(256),
                       0, s, n, (int)d, X, w, maxbits, estat);
    KMD_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(scales_krn, dim3(1), dim3(1), 0, s, n, (int)d, maxbits, estat,
                     scale, nonfinite);
  KMD_LAUNCH_CHECK();
}

template <typename T>
bool fast_fused_supported(int64_t d, int64_t k) {
  if (!std::is_same<T, float>::value) return false;
  return plan_launch<int32_t, true, true>(1, d, k).fn != nullptr &&
         plan_launch<int64_t, true, true>(1, d, k).fn != nullptr &&
         plan_launch<int32_t, true, false>(1, d, k).fn != nullptr &&
         plan_launch<int64_t, true, false>(1, d, k).fn != nullptr;
}

// The driver's step for these float shapes runs the few-centroid kernel.
bool fast_smallk_serves(int64_t d, int64_t k) {
  return d >= 1 && d <= kMaxD && k >= 1 && k <= kSmallK && smallk_enabled() &&
         plan_smallk<int32_t, true>(1, d, k).fn != nullptr;
}

template <typename I>
int fast_fused_blocks(int64_t n, int64_t d, int64_t k) {
  return plan_launch<I, true, true>(n, d, k).blocks;
}

template <typename T, typename I>
void launch_fast_fused_fixed(int64_t n, int64_t d, int64_t k, const T *X,
                             const T *w, const T *C, const T *hn, I *labels,
                             const float *scale, const int *nonfinite,
                             void *partials, int64_t n_partials,
                             hipStream_t s, const int *skip) {
  if constexpr (std::is_same<T, float>::value) {
    if (n <= 0) return;
    const MfmaLaunch L = plan_launch<I, true, true>(n, d, k);
    if (L.fn == nullptr) throw RuntimeError("MFMA fused step: no launch plan");
    if (n_partials < L.blocks)
      throw RuntimeError("fused step needs one partial slot per block");
    FastArgs a = make_args(n, d, k, X, w, C, hn, labels);
    a.vec = vec_ok(X, n, L.s) ? 1 : 0;
    a.scale = scale;
    a.nonfinite = nonfinite;
    a.copies = partials;
    a.n_copies = n_partials;
    a.skip = skip;
    a.n_priv = L.copies;
    run(L, a, s);
  } else {
    (void)skip;
    throw RuntimeError("MFMA fused step is FP32 only");
  }
}

// The reference's fused step (private copies, FP32 atomics): no fixed-point
// scales and so no pre-pass over X; the block sums are FP32 in LDS and reach
// the copies through FP32 atomics, order-dependent like the reference's.
template <typename T, typename I>
void launch_fast_fused(int64_t n, int64_t d, int64_t k, int64_t n_copies,
                       const T *X, const T *w, const T *C, const T *hn,
                       I *labels, T *copies, T *sizes_copies, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    if (n <= 0) return;
    const MfmaLaunch L = plan_launch<I, true, false>(n, d, k);
    if (L.fn == nullptr) throw RuntimeError("MFMA fused step: no launch plan");
    FastArgs a = make_args(n, d, k, X, w, C, hn, labels);
    a.vec = vec_ok(X, n, L.s) ? 1 : 0;
    a.scale = nullptr;
    a.nonfinite = nullptr;
    // window width from the most samples one block adds into one sum:
    // 2^(24 + W) * count < 2^63
    int64_t per_block =
        ceil_div<int64_t>(ceil_div<int64_t>(n, 32 * L.s), L.blocks) * 32 * L.s;
    if (L.smallk) {
      // waves of a block each take every (blocks x waves)-th super-block of 64
      const int64_t waves = L.threads / kWave;
      per_block = ceil_div<int64_t>(ceil_div<int64_t>(n, kWave), L.blocks * waves) *
                  waves * kWave;
    }
    int bits = 0;
    while (bits < 40 && (int64_t(1) << bits) < per_block) ++bits;
    a.window = std::max(1, std::min(30, 38 - bits));
    a.copies = copies;
    a.sizes_copies = sizes_copies;
    a.n_copies = n_copies;
    a.n_priv = L.copies;
    run(L, a, s);
  } else {
    throw RuntimeError("MFMA fused step is FP32 only");
  }
}

template <typename T, typename I>
void launch_fast_assign(int64_t n, int64_t d, int64_t k, const T *X,
                        const T *C, const T *hn, I *labels, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    if (n <= 0) return;
    const MfmaLaunch L = plan_launch<I, false, false>(n, d, k);
    if (L.fn == nullptr) throw RuntimeError("MFMA assignment: no launch plan");
    FastArgs a = make_args(n, d, k, X, nullptr, C, hn, labels);
    a.n_copies = 1;
    run(L, a, s);
  } else {
    throw RuntimeError("MFMA assignment is FP32 only");
  }
}

// ---- FP32 re-score of queued samples (mixed-precision path) ---------------

namespace {
template <typename I>
KernFn pick_rescore(int64_t d) {
  const int ks = ks_bucket(d);
  const bool full = (d == 2 * ks);
#define KMD_KS_CASE(KS)                                                     \
  case KS:                                                                  \
    return full ? &mfma_lloyd_krn<kThreads, KS, kS, true, I, true, true, true> \
                : &mfma_lloyd_krn<kThreads, KS, kS, false, I, true, true, true>;
  switch (ks) {
    KMD_KS_CASE(2)
    KMD_KS_CASE(4)
    KMD_KS_CASE(8)
    KMD_KS_CASE(12)
    KMD_KS_CASE(16)
    KMD_KS_CASE(24)
    KMD_KS_CASE(32)
    default:
      return nullptr;
  }
#undef KMD_KS_CASE
}

template <typename I>
MfmaLaunch plan_rescore(int64_t d, int64_t k) {
  MfmaLaunch L;
  if (d < 1 || d > kMaxD || k < 1) return L;
  L.lds = lds_bytes(d, k, true) + sizeof(int) * (kMaxSegments + 1);
  if (L.lds > kLdsPerCu) return L;
  L.fn = pick_rescore<I>(d);
  if (L.fn == nullptr) return L;
  allow_full_lds(reinterpret_cast<const void *>(L.fn));
  int per_cu = 0;
  KMD_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void *>(L.fn), L.threads, L.lds));
  per_cu = apply_blocks_cap(per_cu);
  if (per_cu < 1) {
    L.fn = nullptr;
    return L;
  }
  const int64_t blocks = (int64_t)device_cu_count() * per_cu;
  L.blocks = (int)(blocks < 1 ? 1 : blocks);
  return L;
}
}  // namespace

template <typename I>
int fast_rescore_blocks(int64_t d, int64_t k, int n_segments) {
  if (n_segments < 1 || n_segments > kMaxSegments) return 0;
  return plan_rescore<I>(d, k).blocks;
}

template <typename I>
void launch_fast_rescore(int64_t n, int64_t d, int64_t k, const float *X,
                         const float *w, const float *C, const float *hn,
                         I *labels, const float *scale, const int *nonfinite,
                         void *partials, int64_t slot0, const int *queue,
                         const int *queue_counts, int n_segments,
                         int64_t seg_cap, int *n_rescored, hipStream_t s) {
  const MfmaLaunch L = plan_rescore<I>(d, k);
  if (L.fn == nullptr) throw RuntimeError("FP32 re-score: no launch plan");
  if (n_segments < 1 || n_segments > kMaxSegments)
    throw RuntimeError("FP32 re-score: segment count out of range");
  FastArgs a = make_args(n, d, k, X, w, C, hn, labels);
  a.vec = 0;
  a.scale = scale;
  a.nonfinite = nonfinite;
  a.copies = partials;
  a.n_copies = L.blocks;
  a.slot0 = slot0;
  a.queue = queue;
  a.queue_counts = queue_counts;
  a.n_segments = n_segments;
  a.seg_cap = seg_cap;
  a.n_rescored = n_rescored;
  run(L, a, s);
}

template int fast_rescore_blocks<int32_t>(int64_t, int64_t, int);
template int fast_rescore_blocks<int64_t>(int64_t, int64_t, int);
template void launch_fast_rescore<int32_t>(int64_t, int64_t, int64_t,
                                           const float *, const float *,
                                           const float *, const float *,
                                           int32_t *, const float *,
                                           const int *, void *, int64_t,
                                           const int *, const int *, int,
                                           int64_t, int *, hipStream_t);
template void launch_fast_rescore<int64_t>(int64_t, int64_t, int64_t,
                                           const float *, const float *,
                                           const float *, const float *,
                                           int64_t *, const float *,
                                           const int *, void *, int64_t,
                                           const int *, const int *, int,
                                           int64_t, int *, hipStream_t);
template int fast_fused_blocks<int32_t>(int64_t, int64_t, int64_t);
template int fast_fused_blocks<int64_t>(int64_t, int64_t, int64_t);
template bool fast_fused_supported<float>(int64_t, int64_t);
template bool fast_fused_supported<double>(int64_t, int64_t);
#define KMD_INST_FAST(T, I)                                                   \
  template void launch_fast_fused<T, I>(int64_t, int64_t, int64_t, int64_t,   \
                                        const T *, const T *, const T *,      \
                                        const T *, I *, T *, T *,             \
                                        hipStream_t);                         \
  template void launch_fast_fused_fixed<T, I>(                                \
      int64_t, int64_t, int64_t, const T *, const T *, const T *, const T *,  \
      I *, const float *, const int *, void *, int64_t, hipStream_t,         \
      const int *);                                                           \
  template void launch_fast_assign<T, I>(int64_t, int64_t, int64_t,           \
                                         const T *, const T *, const T *,     \
                                         I *, hipStream_t);
KMD_INST_FAST(float, int32_t)
KMD_INST_FAST(float, int64_t)
KMD_INST_FAST(double, int32_t)
KMD_INST_FAST(double, int64_t)

}  // namespace kmd
