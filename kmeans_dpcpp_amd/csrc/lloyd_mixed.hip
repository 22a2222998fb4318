// Mixed-precision Lloyd step: BF16 MFMA cross term with a certified margin
// test and exact FP32 re-scoring of undecided samples.
//
// Contraction.  With v_mfma_f32_32x32x16_bf16 (rows = 32 centroids, columns
// = 32 samples, K = 16) each sample/centroid pair gets
//   D = sum_f bf16(x_f) bf16(c_f) - (h1 + h2 + h3)                   (T = 1)
//   D = sum_f (xh ch + xh cl + xl ch) - (h1 + h2 + h3)               (T = 3)
// accumulated in FP32, where xh = bf16(x), xl = bf16(x - xh) (likewise for
// c) and h1 + h2 + h3 is the FP32 half-norm split exactly into three BF16
// terms, so hn is folded into the matrix product (the samples' side of those
// K columns is -1).  The argmin of hn - dot is the argmax of D.
//
// Decision.  Each lane keeps the best and second-best D of its sample (its
// row within the 32-centroid tile packed into the 4 low mantissa bits of
// the value, the tile of the best tracked beside it: one v_and_or per
// candidate and a v_max3 and two v_med3 per pair of candidates).  Let E1 bound the error
// of D against exact arithmetic (bf16 rounding of the operands, FP32
// accumulation, index packing) and E2 the error of the FP32 path's
// pseudo-inertia (FP32 MFMA accumulation and the final subtraction), both
// from ||x||, max ||c|| and max hn.  If best - second > 2 (E1 + E2) the FP32
// path has the same, unique, argmin; the sample is labelled here and its
// w x added to the block's int64 fixed-point sums.  Otherwise its index is
// appended to a queue and the FP32 MFMA kernel re-scores it with exactly the
// FP32 path's arithmetic (same instruction sequence per sample/centroid
// pair), so every label equals the FP32 path's label.  Both kernels add into
// exact int64 partial sums, so which kernel handled a sample does not change
// the sums.
#include "fp32_tile.hpp"
#include "lloyd_fast.hpp"

namespace kmd {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned long long u64;

// threads per block: 16 waves (4 per SIMD) where the registers allow it
// (d <= 32 at one 32-sample column block per wave: <= 128 VGPRs; measured
// faster than 12 waves x 2 column blocks at n = 10 M, d = 32, k = 256)
// (the in-kernel re-score's FP32 registers on top: 12 waves, <= 168 VGPRs)
constexpr int nt_of(int D, bool ink = false) {
  return D <= 32 ? (ink ? 768 : 1024) : 512;
}
constexpr int kS = 1;                  // 32-sample column blocks per wave
constexpr size_t kLdsPerCu = 160 * 1024;
constexpr int kScPad = 66;             // scale slots in LDS (d + 1 <= 65)

template <int D, int T>
struct MixedShape {
  static constexpr int KD = D / 16;                 // k-steps per segment
  static constexpr int KK = (T * D + 3 + 15) / 16;  // k-steps incl. hn
  static constexpr int XH = D / 2;                  // features per lane
};

struct MixedArgs {
  int64_t n;
  int d, k, kpad, idx_bits;
  const float *X, *w;
  const bf16x8 *aop;     // [n_ct][KK][64] A operands
  const float *bound;    // [2]: max ||c||, max hn
  const float *scale;    // [d + 1] fixed-point scales
  void *labels;
  u64 *partials;         // [slot][d*k + k]
  int *queue;            // [block][seg_cap] samples left for the FP32 re-score
  int *queue_counts;     // [block] entries used in each segment
  unsigned *qmask;       // [block][seg_cap] candidate tiles of each queued
                         // sample (k <= 1024), or null: no candidate tracking
  const float *C;        // in-kernel re-score: FP32 centroids [d][k]
  const float *hn;       //   and half norms [k]
  int64_t seg_cap;
  float pinf;            // +inf, passed at run time so that med3(b, v, pinf)
                         // stays one v_med3 (no max pattern, no canonicalise)
  const int *skip;       // when non-null and *skip != 0 the launch does nothing
};

__device__ inline __bf16 to_bf16(float x) { return (__bf16)x; }

// v_max3_f32 on values known to be finite (written out so that no NaN
// canonicalisation is inserted in front of it)
__device__ inline float max3f(float a, float b, float c) {
  float r;
  asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// A operands: element j of lane l for (tile ct, k-step kk) is row
// m = l % 32 (centroid ct*32 + m) at K position p = 16 kk + 8 (l / 32) + j.
template <int D, int T>
__global__ void __launch_bounds__(256)
    prep_aop_krn(int d, int k, int kpad, const float *__restrict__ C,
                 const float *__restrict__ hn, bf16x8 *__restrict__ aop) {
  constexpr int KK = MixedShape<D, T>::KK;
  const int64_t total = (int64_t)(kpad / 32) * KK * 64;
  const int64_t gid = blockIdx.x * 256ll + threadIdx.x;
  if (gid >= total) return;
  const int lane = (int)(gid % 64);
  const int kk = (int)((gid / 64) % KK);
  const int ct = (int)(gid / (64 * KK));
  const int m = lane % 32, half = lane / 32;
  const int c = ct * 32 + m;
  bf16x8 v;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int p = 16 * kk + 8 * half + j;
    float val = 0.f;
    if (p < T * D) {
      const int seg = p / D, f = p % D;
      const float cv = (f < d && c < k) ? C[(int64_t)f * k + c] : 0.f;
      const float hi = (float)to_bf16(cv);
      val = (T == 3 && seg == 1) ? (float)to_bf16(cv - hi) : hi;
    } else if (p < T * D + 3) {
      const int t = p - T * D;
      if (c < k) {
        const float h = hn[c];
        const float h1 = (float)to_bf16(h);
        const float r1 = h - h1;
        const float h2 = (float)to_bf16(r1);
        const float h3 = (float)to_bf16(r1 - h2);
        val = (t == 0) ? h1 : (t == 1 ? h2 : h3);
      } else {
        val = (t == 0) ? 3.0e38f : 0.f;  // padding centroid: never the best
      }
    }
    v[j] = to_bf16(val);
  }
  aop[gid] = v;
}

// bound[0] = max_c ||c|| = sqrt(2 max hn), bound[1] = max_c hn (one block).
__global__ void __launch_bounds__(256)
    bound_krn(int k, const float *__restrict__ hn, float *bound) {
  __shared__ float red[256 / kWave];
  float m = 0.f;
  for (int c = threadIdx.x; c < k; c += 256) m = fmaxf(m, hn[c]);
  for (int o = kWave / 2; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int i = 0; i < 256 / kWave; ++i) t = fmaxf(t, red[i]);
    bound[0] = sqrtf(2.f * t);
    bound[1] = t;
  }
}

// FULL: d == D (no feature beyond d in the lane layout).  NT threads per
// block: one block per CU (the LDS sums fill most of it), so NT sets the
// waves per SIMD that hide each other's load and matrix-core latency.
// INK: undecided samples are re-scored inside this kernel, in the FP32
// path's arithmetic (fp32_tile.hpp over FP32 centroid tiles in LDS): each
// wave keeps its undecided samples in an LDS list and scores 32 of them at
// a time, reading their features again (L2-resident: the wave has just
// read them), then adds them to the block's sums; no queue leaves the
// kernel.  Otherwise the queue segments go to the separate re-score.
template <int D, int T, bool FULL, typename I, int NT, bool INK>
__global__ void __launch_bounds__(NT) mixed_lloyd_krn(MixedArgs a) {
  // an iteration enqueued ahead of a relocation: nothing to compute
  if (a.skip != nullptr && *a.skip != 0) return;
  using Sh = MixedShape<D, T>;
  constexpr int KD = Sh::KD, KK = Sh::KK, XH = Sh::XH;
  constexpr int kWavesPerBlock = NT / kWave;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int64_t n = a.n;
  const int d = a.d, k = a.k, n_ct = a.kpad / 32;
  const int64_t dk = (int64_t)d * k;
  bf16x8 *Al = reinterpret_cast<bf16x8 *>(lds_raw);           // [n_ct*KK*64]
  float *Sc = reinterpret_cast<float *>(Al + (int64_t)n_ct * KK * 64);
  u64 *S64 = reinterpret_cast<u64 *>(Sc + kScPad);             // [d*k + k]
  int &q_used = *reinterpret_cast<int *>(S64 + dk + k);  // queue entries used
  // candidate tracking: each lane's best score per tile, rounded up to bf16,
  // [wave][tile][lane]
  unsigned short *tml = reinterpret_cast<unsigned short *>(
      reinterpret_cast<unsigned char *>(S64 + dk + k) + 16);
  const bool cand = !INK && a.qmask != nullptr;
  const int tid = threadIdx.x;
  // in-kernel re-score: FP32 tiles Cf[((ct KSF + ks) 2 + half) 32 + col] =
  // C[2 ks + half][32 ct + col], -hn (-inf padded), per-wave lists [wave][64]
  constexpr int KSF = D / 2;
  float *Cf = reinterpret_cast<float *>(reinterpret_cast<unsigned char *>(S64 + dk + k) + 16);
  float *Hf = Cf + 2 * KSF * a.kpad;
  int *wq = reinterpret_cast<int *>(Hf + a.kpad);
  for (int64_t i = tid; i < (int64_t)n_ct * KK * 64; i += NT) Al[i] = a.aop[i];
  for (int f = tid; f <= d; f += NT) Sc[f] = a.scale[f];
  for (int64_t i = tid; i < dk + k; i += NT) S64[i] = 0ull;
  if constexpr (INK) {
    for (int i = tid; i < 2 * KSF * a.kpad; i += NT) {
      const int cl = i & 31, hf = (i >> 5) & 1, blk = i >> 6;
      const int ks = blk % KSF, t = blk / KSF;
      const int f = 2 * ks + hf, c = 32 * t + cl;
      Cf[i] = (f < d && c < k) ? a.C[(int64_t)f * k + c] : 0.f;
    }
    for (int c = tid; c < a.kpad; c += NT)
      Hf[c] = (c < k) ? -a.hn[c] : -__builtin_huge_valf();
  }
  if (tid == 0) q_used = 0;
  int *qseg = INK ? nullptr : a.queue + (int64_t)blockIdx.x * a.seg_cap;
  __syncthreads();

  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;
  const int col = lane & 31;
  const int half = lane >> 5;
  // the lane's row r (0..15) of a tile is packed into the low 4 mantissa
  // bits of each value; the tile of the running best is tracked beside it
  constexpr unsigned kRowMask = 15u;
  // error bound E1 + E2 = ||x|| cmax alpha + hnmax beta + pack (|b1| + |b2|)
  const float cmax = a.bound[0], hnmax = a.bound[1];
  // the split of hn into bf16 terms and the bound need a finite, moderate hn
  const bool finite_c = hnmax < 1e36f;
  const float pinf = a.pinf;
  constexpr float u = 1.0f / 256.0f;  // bf16 unit roundoff
  constexpr float acc_err = (float)(T * D + 3 + D + 8) * 5.9604645e-8f;
  const float alpha = ((T == 1) ? (2.f * u + u * u) : (3.2f * u * u)) + acc_err;
  const float beta = acc_err;
  const float pack_err = ldexpf(1.0f, 4 - 23) * 1.01f;
  const float *X = a.X;
  const float *w = a.w;
  I *labels = reinterpret_cast<I *>(a.labels);
  const bool vec = ((reinterpret_cast<uintptr_t>(X) & 7u) == 0) && (n % 2 == 0);

  constexpr int kSuper = 32 * kS;
  const int64_t n_super = ceil_div<int64_t>(n, kSuper);
  // INK: the wave's list of undecided samples and its length (wave-uniform)
  static_assert(!INK || kS == 1, "one 32-sample block per wave and iteration");
  int *wl = wq + wave * kWave;
  int wl_len = 0, wl_done = 0;
  // orders the list's LDS stores and loads across the lanes of the wave
  auto wave_sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  };
  auto rescore_list = [&](int cnt) {
    // lanes l and l ^ 32 take list entry l % 32 (entries cnt.. are padding)
    const int64_t s = col < cnt ? (int64_t)wl[col] : -1;
    float xq[KSF][1];
#pragma unroll
    for (int ks = 0; ks < KSF; ++ks) {
      const int f = 2 * ks + half;
      xq[ks][0] = (s >= 0 && (FULL || f < d)) ? X[(int64_t)f * n + s] : 0.f;
    }
    float bv[1] = {__builtin_huge_valf()};
    int bt[1] = {0}, br[1] = {0};
    for (int ct = 0; ct < n_ct; ++ct) {
      f32x16 acc[1];
      tile_mfma<KSF, 1>(Cf + lane + ct * (KSF * 64), Hf + ct * 32 + 4 * half, xq, acc);
      tile_argmin<1>(ct, acc, bv, bt, br);
    }
    int bi = (bv[0] < __builtin_huge_valf()) ? best_index(bt[0], br[0], half) : 0;
    const float ov = __shfl_xor(bv[0], 32, kWave);
    const int oi = __shfl_xor(bi, 32, kWave);
    if ((ov < bv[0]) || (ov == bv[0] && oi < bi)) bi = oi;
    if (s >= 0) {
      if (half == 0) labels[s] = (I)bi;
      const float ws = w[s];
      u64 *row = S64 + (int64_t)bi * (d + 1);
#pragma unroll
      for (int ks = 0; ks < KSF; ++ks) {
        const int f = 2 * ks + half;
        if (FULL || f < d)
          atomicAdd(row + f, (u64)(long long)__float2int_rn((xq[ks][0] * ws) * Sc[f]));
      }
      if (half == 0)
        atomicAdd(row + d, (u64)(long long)__float2int_rn(ws * Sc[d]));
    }
    wl_done += cnt;
  };
  // lane features f(i) = 16 (i / 8) + 8 half + i % 8
  auto load = [&](int64_t s0, float (&dst)[XH][kS]) {
    const bool whole = vec && (s0 + kS <= n);
#pragma unroll
    for (int i = 0; i < XH; ++i) {
      const int f = 16 * (i / 8) + 8 * half + (i % 8);
      if (FULL || f < d) {
        const float *row = X + (int64_t)f * n + s0;
        if (whole) {
          if constexpr (kS == 2) {
            const float2 v = *reinterpret_cast<const float2 *>(row);
            dst[i][0] = v.x;
            dst[i][1] = v.y;
          } else {
            dst[i][0] = row[0];
          }
        } else {
#pragma unroll
          for (int q = 0; q < kS; ++q) dst[i][q] = (s0 + q < n) ? row[q] : 0.f;
        }
      } else {
#pragma unroll
        for (int q = 0; q < kS; ++q) dst[i][q] = 0.f;
      }
    }
  };
  const int64_t sb_stride = (int64_t)gridDim.x * kWavesPerBlock;
  for (int64_t sb = (int64_t)blockIdx.x * kWavesPerBlock + wave; sb < n_super;
       sb += sb_stride) {
    const int64_t s0 = sb * kSuper + (int64_t)kS * col;
    // x is held only until it is converted (registers go to the other
    // waves); the update reads it again, from L2
    float xr[XH][kS];
    load(s0, xr);
    float xnorm[kS];
#pragma unroll
    for (int q = 0; q < kS; ++q) {
      float ss = 0.f;
#pragma unroll
      for (int i = 0; i < XH; ++i) ss = fmaf(xr[i][q], xr[i][q], ss);
      ss += __shfl_xor(ss, 32, kWave);
      xnorm[q] = sqrtf(ss) * 1.0001f;
    }
    // B operands (samples): bf16 high parts, low parts (T = 3) and the
    // -1 entries that pick up the half-norm columns
    bf16x8 bhi[KD][kS];
    bf16x8 blo[(T == 3) ? KD : 1][kS];
#pragma unroll
    for (int kd = 0; kd < KD; ++kd)
#pragma unroll
      for (int q = 0; q < kS; ++q)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const float x = xr[8 * kd + j][q];
          const __bf16 h = to_bf16(x);
          bhi[kd][q][j] = h;
          if constexpr (T == 3) blo[kd][q][j] = to_bf16(x - (float)h);
        }
    bf16x8 bh;
#pragma unroll
    for (int j = 0; j < 8; ++j)
      bh[j] = to_bf16((half == 0 && j < 3) ? -1.f : 0.f);

    float b1[kS], b2[kS];
    int bt[kS];
#pragma unroll
    for (int q = 0; q < kS; ++q) {
      b1[q] = -__builtin_huge_valf();
      b2[q] = -__builtin_huge_valf();
      bt[q] = 0;
    }
    // products of tile ct into acc (K steps: x high parts, then for T = 3
    // the two cross terms, then the half-norm columns)
    auto tile_mma = [&](int ct, f32x16(&acc)[kS]) {
      const bf16x8 *At = Al + (int64_t)ct * KK * 64 + lane;
#pragma unroll
      for (int kk = 0; kk < KK; ++kk) {
        const bf16x8 av = At[kk * 64];
#pragma unroll
        for (int q = 0; q < kS; ++q) {
          bf16x8 bv;
          if (kk < KD)
            bv = bhi[kk][q];
          else if (T == 3 && kk < 2 * KD)
            bv = bhi[(kk - KD) % KD][q];
          else if (T == 3 && kk < 3 * KD)
            bv = blo[(kk - 2 * KD) % KD][q];
          else
            bv = bh;
          if (kk == 0) {
            const f32x16 z = {};
            acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, z, 0, 0, 0);
          } else {
            acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[q], 0,
                                                              0, 0);
          }
        }
      }
    };
    // top-2 update with the 16 rows of tile ct this lane holds (row r is
    // centroid 32 ct + 8 (r / 4) + 4 half + (r % 4)), r packed into the low
    // bits (one v_and_or with an inline constant per value).
    // Two candidates u, v at a time: top-2 of {b1, b2, u, v} (b1 >= b2) is
    // b1' = max3(b1, u, v), b2' = max(b2, med3(b1, u, v)).  med3(x, y, +inf)
    // is max(x, y) without the NaN canonicalisation of a max; the values are
    // finite here or the sample is re-scored.  The best's tile is the last
    // tile in which b1 rose (a tie with an earlier tile's value leaves a gap
    // of 0: such a sample is re-scored).
    auto tile_track = [&](int ct, const f32x16(&acc)[kS]) {
      float tmax[kS], b1_in[kS];
#pragma unroll
      for (int q = 0; q < kS; ++q) {
        tmax[q] = -__builtin_huge_valf();
        b1_in[q] = b1[q];
      }
#pragma unroll
      for (int r = 0; r < 16; r += 2) {
#pragma unroll
        for (int q = 0; q < kS; ++q) {
          const float u = __uint_as_float(
              (__float_as_uint(acc[q][r]) & ~kRowMask) | (unsigned)r);
          const float v = __uint_as_float(
              (__float_as_uint(acc[q][r + 1]) & ~kRowMask) | (unsigned)(r + 1));
          const float m = __builtin_amdgcn_fmed3f(b1[q], u, v);
          b1[q] = max3f(b1[q], u, v);
          b2[q] = __builtin_amdgcn_fmed3f(b2[q], m, pinf);
          if (cand) tmax[q] = max3f(tmax[q], u, v);
        }
      }
#pragma unroll
      for (int q = 0; q < kS; ++q) bt[q] = (b1[q] != b1_in[q]) ? ct : bt[q];
      // the tile's best, rounded toward +inf to bf16 (never below it)
      if (cand) {
        static_assert(kS == 1, "one tile maximum per lane and tile");
        const unsigned b = __float_as_uint(tmax[0]);
        const unsigned up = (b & 0x80000000u) ? (b >> 16) : ((b + 0xffffu) >> 16);
        tml[(wave * n_ct + ct) * kWave + lane] = (unsigned short)up;
      }
    };
    if constexpr (D <= 16) {
      // the products of the next tile run on the matrix core while the
      // vector unit ranks the current one
      f32x16 acc0[kS], acc1[kS];
      tile_mma(0, acc0);
      for (int ct = 0; ct < n_ct; ct += 2) {
        const bool has1 = ct + 1 < n_ct;
        if (has1) tile_mma(ct + 1, acc1);
        tile_track(ct, acc0);
        if (has1) {
          if (ct + 2 < n_ct) tile_mma(ct + 2, acc0);
          tile_track(ct + 1, acc1);
        }
      }
    } else {
      // wider samples leave no registers for a second accumulator set
      for (int ct = 0; ct < n_ct; ++ct) {
        f32x16 acc[kS];
        tile_mma(ct, acc);
        tile_track(ct, acc);
      }
    }
    int lab[kS];
#pragma unroll
    for (int q = 0; q < kS; ++q) {
      lab[q] = best_index(bt[q], (int)(__float_as_uint(b1[q]) & kRowMask), half);
      const float o1 = __shfl_xor(b1[q], 32, kWave);
      const float o2 = __shfl_xor(b2[q], 32, kWave);
      const int ol = __shfl_xor(lab[q], 32, kWave);
      const float n2 = fmaxf(fminf(b1[q], o1), fmaxf(b2[q], o2));
      lab[q] = (o1 > b1[q]) ? ol : lab[q];
      b1[q] = fmaxf(b1[q], o1);
      b2[q] = n2;
    }
    bool amb[kS];
#pragma unroll
    for (int q = 0; q < kS; ++q) {
      const float e = xnorm[q] * cmax * alpha + hnmax * beta +
                      pack_err * (fabsf(b1[q]) + fabsf(b2[q]));
      const float gap = b1[q] - b2[q];
      // far from overflow |D| <= ||x|| ||c|| (1 + 2^-8)^2 + hn stays finite
      amb[q] = !(gap > 2.f * e) || !(xnorm[q] * cmax < 1e36f) || !finite_c ||
               lab[q] >= k;
    }
    // labels of decided samples; queue the others for the FP32 re-score
#pragma unroll
    for (int q = 0; q < kS; ++q) {
      const int64_t s = s0 + q;
      const bool valid = s < n;
      if (half == 0 && valid && !amb[q]) labels[s] = (I)lab[q];
      const bool enq = (half == 0) && valid && amb[q];
      // candidate tiles against the final best: a centroid can be the FP32
      // argmin only if its score is within 2 (E1 + E2) of the best; with the
      // packing error of both values this is 2 e_part + 6 pack M
      unsigned tiles_any = 0u;
      if (cand) {
        const float big = 1.01f * (xnorm[q] * cmax + hnmax);
        const float thr = b1[q] - (2.f * (xnorm[q] * cmax * alpha + hnmax * beta) +
                                   6.f * pack_err * big);
        unsigned bits = 0u;
        for (int t = 0; t < n_ct; ++t) {
          const float v = __uint_as_float(
              (unsigned)tml[(wave * n_ct + t) * kWave + lane] << 16);
          if (v >= thr) bits |= 1u << t;
        }
        tiles_any = bits | (unsigned)__shfl_xor((int)bits, 32, kWave);
      }
      const uint64_t mask = __ballot(enq);
      if (INK) {
        // append to the wave's list (< 32 entries before, <= 32 added)
        if (enq) wl[wl_len + __popcll(mask & ((1ull << lane) - 1ull))] = (int)s;
        wl_len += __popcll(mask);
      } else if (mask != 0ull) {
        const int cnt = __popcll(mask);
        const int leader = __ffsll((long long)mask) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(&q_used, cnt);
        base = __shfl(base, leader, kWave);
        if (enq) {
          const int rank = __popcll(mask & ((1ull << lane) - 1ull));
          qseg[base + rank] = (int)s;
          if (cand) a.qmask[(int64_t)blockIdx.x * a.seg_cap + base + rank] = tiles_any;
        }
      }
    }
    // int64 fixed-point update, sums centroid-major in LDS: S64[l (d+1) + f]
    // (f = d: the count), so each lane's 16 features are immediate offsets
    // from one address.  Undecided and out-of-range samples add zeros
    // (weight 0), which keeps the wave convergent.
    load(s0, xr);
#pragma unroll
    for (int q = 0; q < kS; ++q) {
      const int64_t s = s0 + q;
      const float ws = (s < n && !amb[q]) ? w[s] : 0.f;
      const int l = min(lab[q], k - 1);
      u64 *row = S64 + (int64_t)l * (d + 1);
#pragma unroll
      for (int i = 0; i < XH; ++i) {
        const int f = 16 * (i / 8) + 8 * half + (i % 8);
        if (FULL || f < d) {
          const int qv = __float2int_rn((xr[i][q] * ws) * Sc[f]);
          atomicAdd(row + f, (u64)(long long)qv);
        }
      }
      if (half == 0)
        atomicAdd(row + d, (u64)(long long)__float2int_rn(ws * Sc[d]));
    }
    if (INK && wl_len >= 32) {
      wave_sync();
      rescore_list(32);
      // move the rest (< 32 entries) to the front
      const int rest = wl_len - 32;
      const int v = lane < rest ? wl[32 + lane] : 0;
      wave_sync();
      if (lane < rest) wl[lane] = v;
      wave_sync();
      wl_len = rest;
    }
  }
  if (INK) {
    if (wl_len > 0) {
      wave_sync();
      rescore_list(wl_len);
    }
    if (lane == 0 && wl_done > 0) atomicAdd(&q_used, wl_done);
  }
  __syncthreads();
  // partial slot layout: [f][c] sums, then [c] counts
  u64 *cp = a.partials + (int64_t)blockIdx.x * (dk + k);
  for (int64_t i = tid; i < dk + k; i += NT) {
    const int64_t f = i / k, c = i - f * k;
    cp[i] = S64[c * (d + 1) + f];
  }
  if (tid == 0) a.queue_counts[blockIdx.x] = q_used;
}

inline int d_bucket(int64_t d) {
  if (d <= 16) return 16;
  if (d <= 32) return 32;
  if (d <= 48) return 48;
  return 64;
}

inline int kk_of(int D, int T) { return (T * D + 3 + 15) / 16; }

// cand: with the per-tile maxima of every lane (candidate tiles, k <= 1024);
// ink: with the FP32 tiles, -hn and the per-wave lists of the in-kernel
// re-score
inline size_t mixed_lds_bytes(int64_t d, int64_t k, int T, bool cand = false,
                              bool ink = false) {
  const int D = d_bucket(d);
  const int64_t kpad = ceil_div<int64_t>(k, 32) * 32;
  const size_t tml = cand ? sizeof(unsigned short) * (size_t)(nt_of(D, false) / kWave) *
                                (size_t)(kpad / 32) * kWave
                          : 0;
  const size_t inks = ink ? sizeof(float) * (size_t)(D + 1) * kpad +
                                sizeof(int) * (size_t)nt_of(D, true)
                          : 0;
  return (size_t)(kpad / 32) * kk_of(D, T) * 64 * 16 + sizeof(float) * kScPad +
         sizeof(u64) * (size_t)(d + 1) * k + 16 + std::max(tml, inks);
}

using MixedFn = void (*)(MixedArgs);

template <typename I, bool INK>
MixedFn pick_mixed_ink(int64_t d, int T) {
  const int D = d_bucket(d);
#define KMD_MIXED_CASE(DD)                                                  \
  if (D == DD) {                                                            \
    constexpr int NT = nt_of(DD, INK);                                      \
    if (d == DD)                                                            \
      return T == 3 ? &mixed_lloyd_krn<DD, 3, true, I, NT, INK>             \
                    : &mixed_lloyd_krn<DD, 1, true, I, NT, INK>;            \
    return T == 3 ? &mixed_lloyd_krn<DD, 3, false, I, NT, INK>              \
                  : &mixed_lloyd_krn<DD, 1, false, I, NT, INK>;             \
  }
  KMD_MIXED_CASE(16)
  KMD_MIXED_CASE(32)
  KMD_MIXED_CASE(48)
  KMD_MIXED_CASE(64)
#undef KMD_MIXED_CASE
  return nullptr;
}

template <typename I>
MixedFn pick_mixed(int64_t d, int T, bool ink = false) {
  return ink ? pick_mixed_ink<I, true>(d, T) : pick_mixed_ink<I, false>(d, T);
}

}  // namespace

int mixed_idx_bits(int64_t k) {
  const int64_t kpad = ceil_div<int64_t>(k, 32) * 32;
  int b = 1;
  while ((int64_t(1) << b) < kpad) ++b;
  return b;
}

size_t mixed_aop_bytes(int64_t d, int64_t k, int T) {
  const int D = d_bucket(d);
  const int64_t kpad = ceil_div<int64_t>(k, 32) * 32;
  return (size_t)(kpad / 32) * kk_of(D, T) * 64 * 16;
}

int64_t mixed_seg_cap(int64_t n, int64_t d, int blocks) {
  const int nt = nt_of(d_bucket(d));
  const int64_t n_super = ceil_div<int64_t>(n > 0 ? n : 1, 32 * kS);
  const int64_t waves = (int64_t)blocks * (nt / kWave);
  return ceil_div<int64_t>(n_super, waves) * (nt / kWave) * 32 * kS;
}

bool mixed_cand_fits(int64_t d, int64_t k, int T) {
  return k <= 1024 && mixed_lds_bytes(d, k, T, true) <= kLdsPerCu;
}

bool mixed_ink_fits(int64_t d, int64_t k, int T) {
  // d <= 48: the FP32 re-score's registers fit beside the BF16 kernel's
  // without spilling (not at d = 64)
  return d <= 48 && mixed_lds_bytes(d, k, T, false, true) <= kLdsPerCu;
}

template <typename I>
int mixed_blocks(int64_t n, int64_t d, int64_t k, int T, bool cand, bool ink) {
  if (d < 1 || d > 64 || k < 1 || n >= (int64_t(1) << 31)) return 0;
  if (mixed_idx_bits(k) > 12) return 0;  // index packing: <= 4096 centroids
  if (cand && ink) return 0;
  if (cand && !mixed_cand_fits(d, k, T)) return 0;
  const size_t lds = mixed_lds_bytes(d, k, T, cand, ink);
  if (lds > kLdsPerCu) return 0;
  MixedFn fn = pick_mixed<I>(d, T, ink);
  if (fn == nullptr) return 0;
  static thread_local MixedFn attr_done[64] = {};
  bool seen = false;
  for (MixedFn f : attr_done) seen |= (f == fn);
  if (!seen) {
    KMD_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)kLdsPerCu));
    for (MixedFn &f : attr_done)
      if (f == nullptr) {
        f = fn;
        break;
      }
  }
  int per_cu = 0;
  KMD_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void *>(fn), nt_of(d_bucket(d), ink), lds));
  per_cu = apply_blocks_cap(per_cu);
  if (per_cu < 1) return 0;
  const int64_t need =
      ceil_div<int64_t>(ceil_div<int64_t>(n > 0 ? n : 1, 32 * kS),
                        nt_of(d_bucket(d), ink) / kWave);
  int64_t blocks = (int64_t)device_cu_count() * per_cu;
  if (blocks > need) blocks = need;
  if (blocks > 1024) blocks = 1024;  // queue segments the re-score reads
  return (int)(blocks < 1 ? 1 : blocks);
}

void launch_mixed_prep(int64_t d, int64_t k, int T, const float *C,
                       const float *hn, void *aop, float *bound,
                       hipStream_t s) {
  const int D = d_bucket(d);
  const int kpad = (int)(ceil_div<int64_t>(k, 32) * 32);
  const int64_t total = (int64_t)(kpad / 32) * kk_of(D, T) * 64;
  const dim3 grid((unsigned)ceil_div<int64_t>(total, 256));
#define KMD_PREP_CASE(DD)                                                     \
  if (D == DD) {                                                              \
    if (T == 3)                                                               \
      hipLaunchKernelGGL((prep_aop_krn<DD, 3>), grid, dim3(256), 0, s, (int)d, \
                         (int)k, kpad, C, hn, (bf16x8 *)aop);                 \
    else                                                                      \
      hipLaunchKernelGGL((prep_aop_krn<DD, 1>), grid, dim3(256), 0, s, (int)d, \
                         (int)k, kpad, C, hn, (bf16x8 *)aop);                 \
  }
  KMD_PREP_CASE(16)
  KMD_PREP_CASE(32)
  KMD_PREP_CASE(48)
  KMD_PREP_CASE(64)
#undef KMD_PREP_CASE
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(bound_krn, dim3(1), dim3(256), 0, s, (int)k, hn, bound);
  KMD_LAUNCH_CHECK();
}

namespace {
// total of the per-block counts of the in-kernel re-score (one wave)
__global__ void count_total_krn(int blocks, const int *counts, int *total,
                                const int *skip) {
  if (skip != nullptr && *skip != 0) return;
  int t = 0;
  for (int b = threadIdx.x; b < blocks; b += kWave) t += counts[b];
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) t += __shfl_xor(t, o, kWave);
  if (threadIdx.x == 0) total[0] = t;
}
}  // namespace

template <typename I>
void launch_mixed_step(int64_t n, int64_t d, int64_t k, int T, int blocks,
                       const float *X, const float *w, const void *aop,
                       const float *bound, const float *scale, I *labels,
                       void *partials, int *queue, int *queue_counts,
                       unsigned *qmask, hipStream_t s, const int *skip,
                       const float *C, const float *hn, int *rescored_total) {
  const bool ink = C != nullptr;
  if (ink && (qmask != nullptr || hn == nullptr || rescored_total == nullptr))
    throw RuntimeError("mixed-precision step: in-kernel re-score arguments");
  MixedFn fn = pick_mixed<I>(d, T, ink);
  if (fn == nullptr || blocks < 1)
    throw RuntimeError("mixed-precision step: unsupported shape");
  MixedArgs a{};
  a.n = n;
  a.d = (int)d;
  a.k = (int)k;
  a.kpad = (int)(ceil_div<int64_t>(k, 32) * 32);
  a.idx_bits = mixed_idx_bits(k);
  a.X = X;
  a.w = w;
  a.aop = reinterpret_cast<const bf16x8 *>(aop);
  a.bound = bound;
  a.scale = scale;
  a.labels = labels;
  a.partials = reinterpret_cast<u64 *>(partials);
  a.queue = queue;
  a.queue_counts = queue_counts;
  a.qmask = qmask;
  a.seg_cap = mixed_seg_cap(n, d, blocks);
  a.pinf = std::numeric_limits<float>::infinity();
  a.skip = skip;
  a.C = C;
  a.hn = hn;
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(nt_of(d_bucket(d), ink)),
                     mixed_lds_bytes(d, k, T, qmask != nullptr, ink), s,
                     a);
  KMD_LAUNCH_CHECK();
  if (ink) {
    hipLaunchKernelGGL(count_total_krn, dim3(1), dim3(kWave), 0, s, blocks,
                       (const int *)queue_counts, rescored_total, skip);
    KMD_LAUNCH_CHECK();
  }
}

namespace {
// int64 fixed-point update of the queued (re-scored) samples: block b takes
// segments b, b + grid, ... of the queue, adds rint((x w) scale_f) of each
// sample (x from the sample-major copy X_r) under its label into LDS, and
// writes its partial slot [d k + k] (the FP32 kernels' element arithmetic,
// so the integers are the ones the FP32 path adds for these samples).
template <typename I>
__global__ void __launch_bounds__(256)
    queue_update_krn(int d, int k, int n_segments, int64_t seg_cap,
                     const int *__restrict__ queue, const int *__restrict__ queue_counts,
                     const float *__restrict__ X_r, const float *__restrict__ w,
                     const I *__restrict__ labels, const float *__restrict__ scale,
                     u64 *partials, const int *skip) {
  if (skip != nullptr && *skip != 0) return;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  u64 *S64 = reinterpret_cast<u64 *>(lds_raw);   // [d k + k]
  float *Sc = reinterpret_cast<float *>(S64 + (int64_t)(d + 1) * k);
  const int64_t dk = (int64_t)d * k;
  for (int64_t i = threadIdx.x; i < dk + k; i += 256) S64[i] = 0ull;
  for (int f = threadIdx.x; f <= d; f += 256) Sc[f] = scale[f];
  __syncthreads();
  for (int seg = blockIdx.x; seg < n_segments; seg += gridDim.x) {
    const int m = queue_counts[seg];
    const int *qs = queue + (int64_t)seg * seg_cap;
    for (int j = threadIdx.x; j < m; j += 256) {
      const int64_t s = qs[j];
      const int l = (int)labels[s];
      const float ws = w[s];
      const float *row = X_r + s * d;
      for (int f = 0; f < d; ++f) {
        const int qv = __float2int_rn((row[f] * ws) * Sc[f]);
        atomicAdd(&S64[(int64_t)f * k + l], (u64)(long long)qv);
      }
      atomicAdd(&S64[dk + l], (u64)(long long)__float2int_rn(ws * Sc[d]));
    }
  }
  __syncthreads();
  u64 *cp = partials + (int64_t)blockIdx.x * (dk + k);
  for (int64_t i = threadIdx.x; i < dk + k; i += 256) cp[i] = S64[i];
}
}  // namespace

template <typename I>
void launch_queue_update(int64_t d, int64_t k, int blocks, int n_segments,
                         int64_t seg_cap, const int *queue, const int *queue_counts,
                         const float *X_r, const float *w, const I *labels,
                         const float *scale, void *partials, hipStream_t s,
                         const int *skip) {
  const size_t lds = sizeof(u64) * (size_t)(d + 1) * k + sizeof(float) * (d + 1);
  require(lds <= kLdsPerCu, "queue update: sums beyond LDS");
  if (lds > 64 * 1024) allow_full_lds(reinterpret_cast<const void *>(&queue_update_krn<I>));
  hipLaunchKernelGGL(queue_update_krn<I>, dim3((unsigned)std::max(1, blocks)), dim3(256),
                     lds, s, (int)d, (int)k, n_segments, seg_cap, queue, queue_counts,
                     X_r, w, labels, scale, reinterpret_cast<u64 *>(partials), skip);
  KMD_LAUNCH_CHECK();
}

template int mixed_blocks<int32_t>(int64_t, int64_t, int64_t, int, bool, bool);
template int mixed_blocks<int64_t>(int64_t, int64_t, int64_t, int, bool, bool);
template void launch_mixed_step<int32_t>(int64_t, int64_t, int64_t, int, int,
                                         const float *, const float *,
                                         const void *, const float *,
                                         const float *, int32_t *, void *,
                                         int *, int *, unsigned *, hipStream_t,
                                         const int *, const float *, const float *,
                                         int *);
template void launch_mixed_step<int64_t>(int64_t, int64_t, int64_t, int, int,
                                         const float *, const float *,
                                         const void *, const float *,
                                         const float *, int64_t *, void *,
                                         int *, int *, unsigned *, hipStream_t,
                                         const int *, const float *, const float *,
                                         int *);
template void launch_queue_update<int32_t>(int64_t, int64_t, int, int, int64_t,
                                           const int *, const int *, const float *,
                                           const float *, const int32_t *,
                                           const float *, void *, hipStream_t,
                                           const int *);
template void launch_queue_update<int64_t>(int64_t, int64_t, int, int, int64_t,
                                           const int *, const int *, const float *,
                                           const float *, const int64_t *,
                                           const float *, void *, hipStream_t,
                                           const int *);

}  // namespace kmd
