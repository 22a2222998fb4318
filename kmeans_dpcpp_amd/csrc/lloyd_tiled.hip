// Tiled Lloyd path: shapes the LDS-resident kernels cannot hold (d up to 128,
// k up to 32768), e.g. n = 2M, d = 128, k = 1024.
//
// Assignment.  Samples stay in registers for a whole pass over the
// centroids; the centroids stream through LDS in chunks of 32-centroid tiles
// that the 8 waves of a block share (every block walks the same chunks, so
// they are L2 hits after the first block).  Blocks step through rounds of
// 8 super-blocks in lock step, so every barrier is reached by all waves.
//   FP32: v_mfma_f32_32x32x2f32, the FP32 path's arithmetic (products
//         accumulated in FP32 over ascending k-steps, argmin of hn - dot with
//         strict < and lowest index on ties).  A gather mode re-scores the
//         samples listed in queue segments, reading them from a sample-major
//         copy of X.
//   BF16: v_mfma_f32_32x32x16_bf16 with the half norm folded into the
//         product and a certified best / second-best margin test (the scheme
//         of lloyd_mixed.hip); undecided samples go to per-block queue
//         segments for the FP32 gather re-score, so labels equal the FP32
//         path's.
//
// Update.  With d * k beyond LDS, per-block sums do not fit, so the update is
// its own pass: labels are counted, the cluster totals scanned, sample
// indices scattered into cluster order, and each (cluster, chunk of 256
// members) work item sums w x over its members from the sample-major copy
// X_r (one contiguous row per sample) as int64 fixed point (the scales of
// launch_fixed_point_scales), then adds into the cluster's int64 sums with
// one global atomic per feature.  Integer addition is exact and
// order-independent, so the sums equal those of the fused kernels bit for
// bit.
#include <algorithm>

#include "lloyd_fast.hpp"

namespace kmd {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned long long u64;

constexpr int kTNT = 512;                 // threads per block (8 waves)
constexpr int kTW = kTNT / kWave;         // waves per block
constexpr int kChunkBytes = 52 * 1024;    // LDS bytes of one centroid chunk
                                          // (two of them: double buffered)
constexpr int kMaxSegments = 1024;        // queue segments the re-score reads
constexpr int kMaxTiledD = 128;
constexpr int kMembers = 256;             // members per update work item
constexpr size_t kLdsPerCu = 160 * 1024;

// Copies `bytes` from global `src` (16-byte aligned, readable for 1 KB past
// `bytes`) to LDS `dst` with global_load_lds_dwordx4: wave w moves the 1 KB
// pieces w, w + kTW, ... (64 lanes x 16 B, lane l's bytes landing at
// piece * 1 KB + 16 l).  Does not wait: the caller waits (vmcnt) and
// synchronises before reading `dst`.
__device__ inline void lds_dma(const unsigned char *src, unsigned char *dst,
                               int bytes, int wave, int lane) {
  const int pieces = (bytes + 1023) >> 10;
  for (int p = wave; p < pieces; p += kTW)
    __builtin_amdgcn_global_load_lds(
        (const void *)(src + (size_t)p * 1024 + lane * 16),
        (__attribute__((address_space(3))) void *)(dst + (size_t)p * 1024), 16, 0,
        0);
}

__device__ inline void wait_dma_and_sync() {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
}

// ---- FP32 assignment --------------------------------------------------------

// FP32 k-steps (2 features each), bucketed.
inline int ks_of(int64_t d) {
  const int64_t ks = (d + 1) / 2;
  if (ks <= 16) return 16;
  if (ks <= 32) return 32;
  return 64;
}

template <int KS>
constexpr int tile_floats() {  // A values + the tile's 32 half norms
  return 2 * KS * 32 + 32;
}
template <int KS>
constexpr int fp32_tiles_per_chunk() {
  return (kChunkBytes / (4 * tile_floats<KS>())) < 1
             ? 1
             : kChunkBytes / (4 * tile_floats<KS>());
}
// samples (32-column MFMA blocks) per wave: the registers hold KS * S values
template <int KS>
constexpr int fp32_s() {
  return KS <= 32 ? 2 : 1;
}

struct TiledArgs {
  int64_t n;
  int d, k, n_ct;
  const float *X;      // X_t (d, n); gather mode: X_r (n, d)
  const float *tiles;  // [n_ct][tile_floats] in the LDS layout (+1 KB pad)
  void *labels;
  const int *queue, *queue_counts;  // gather mode
  int n_segments;
  int64_t seg_cap;
  int *n_rescored;
  const int *only_if;  // gather mode: run only when *only_if != 0 (or null)
  const int *skip;     // when non-null and *skip != 0 the launch does nothing
};

// Tile layout in LDS: element (ks, half, col) = C[2 ks + half][tile*32 + col]
// at ((ks / 4) * 2 + half) * 128 + col * 4 + ks % 4, so each lane reads four
// k-steps with one 16-byte load; then the 32 half norms (+inf padded).
template <int KS, typename I, bool GATHER>
__global__ void __launch_bounds__(kTNT) tiled_fp32_krn(TiledArgs a) {
  // an iteration enqueued ahead of a relocation: nothing to compute
  if (a.skip != nullptr && *a.skip != 0) return;
  constexpr int S = fp32_s<KS>();
  constexpr int TF = tile_floats<KS>();
  constexpr int CPC = fp32_tiles_per_chunk<KS>();
  constexpr int kBuf = ((CPC * TF * 4 + 1023) / 1024) * 1024;  // bytes per buffer
  extern __shared__ __align__(16) unsigned char lds_raw[];
  int *seg_off = reinterpret_cast<int *>(lds_raw + 2 * kBuf);  // [n_segments + 1]
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1), wave = tid / kWave;
  const int col = lane & 31, half = lane >> 5;
  const int64_t n = a.n;
  const int d = a.d, n_ct = a.n_ct;
  I *labels = reinterpret_cast<I *>(a.labels);

  if constexpr (GATHER) {
    if (a.only_if != nullptr && a.only_if[0] == 0) return;  // block-uniform
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
    __syncthreads();
  }
  const int64_t n_items = GATHER ? (int64_t)seg_off[a.n_segments] : n;
  constexpr int kSuper = 32 * S;
  const int64_t n_super = ceil_div<int64_t>(n_items, kSuper);
  const int64_t per_round = (int64_t)gridDim.x * kTW;
  const int64_t rounds = ceil_div<int64_t>(n_super, per_round);
  const bool vec = !GATHER && S == 2 &&
                   ((reinterpret_cast<uintptr_t>(a.X) & 7u) == 0) && (n % 2 == 0);

  // centroid chunks stream through two LDS buffers: chunk g + 1 is copied
  // while chunk g is used (the sequence repeats every round)
  const int n_chunks = (n_ct + CPC - 1) / CPC;
  const int64_t total_chunks = rounds * n_chunks;
  const unsigned char *tiles = reinterpret_cast<const unsigned char *>(a.tiles);
  auto chunk_bytes = [&](int c) { return min(CPC, n_ct - c * CPC) * TF * 4; };
  if (total_chunks > 0) lds_dma(tiles, lds_raw, chunk_bytes(0), wave, lane);
  int64_t g = 0;
  for (int64_t r = 0; r < rounds; ++r) {
    const int64_t sb = r * per_round + (int64_t)blockIdx.x * kTW + wave;
    const bool active = sb < n_super;
    const int64_t p0 = sb * kSuper + (int64_t)S * col;
    int64_t sidx[S];
    bool valid[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
      valid[q] = active && (p0 + q < n_items);
      sidx[q] = p0 + q;
      if (GATHER && valid[q]) {
        const int p = (int)(p0 + q);
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
    float xr[KS][S];
    if (GATHER) {
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int f = 2 * ks + half;
#pragma unroll
        for (int q = 0; q < S; ++q)
          xr[ks][q] = (valid[q] && f < d) ? a.X[sidx[q] * d + f] : 0.f;
      }
    } else {
      const bool whole = vec && active && (p0 + S <= n);
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int f = 2 * ks + half;
        const float *row = a.X + (int64_t)f * n + p0;
        if (f < d && whole) {
          if constexpr (S == 2) {
            const float2 v = *reinterpret_cast<const float2 *>(row);
            xr[ks][0] = v.x;
            xr[ks][1] = v.y;
          }
        } else {
#pragma unroll
          for (int q = 0; q < S; ++q)
            xr[ks][q] = (f < d && valid[q]) ? row[q] : 0.f;
        }
      }
    }
    float bestv[S];
    int besti[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
      bestv[q] = __builtin_huge_valf();
      besti[q] = 0;
    }
    for (int ci = 0; ci < n_chunks; ++ci, ++g) {
      const int c0 = ci * CPC;
      const int cnt = min(CPC, n_ct - c0);
      wait_dma_and_sync();
      if (g + 1 < total_chunks) {
        const int cn = (int)((g + 1) % n_chunks);
        lds_dma(tiles + (size_t)cn * CPC * TF * 4, lds_raw + ((g + 1) & 1) * kBuf,
                chunk_bytes(cn), wave, lane);
      }
      const float *Cl = reinterpret_cast<const float *>(lds_raw + (g & 1) * kBuf);
      if (active) {
        for (int t = 0; t < cnt; ++t) {
          const float *At = Cl + t * TF;
          // rows of this lane: m = 8 (r / 4) + 4 half + (r % 4), ascending;
          // the chain starts from -hn (the tile stores it), so -acc = hn - c.x
          const float *Ht = At + 2 * KS * 32;
          f32x16 h;
#pragma unroll
          for (int gg = 0; gg < 4; ++gg) {
            const f32x4 h4 = *reinterpret_cast<const f32x4 *>(Ht + 8 * gg + 4 * half);
#pragma unroll
            for (int e = 0; e < 4; ++e) h[4 * gg + e] = h4[e];
          }
          f32x16 acc[S];
#pragma unroll
          for (int k4 = 0; k4 < KS / 4; ++k4) {
            const f32x4 av =
                *reinterpret_cast<const f32x4 *>(At + (k4 * 2 + half) * 128 + col * 4);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
#pragma unroll
              for (int q = 0; q < S; ++q) {
                if (k4 == 0 && j == 0) {
                  acc[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[j], xr[0][q], h,
                                                                0, 0, 0);
                } else {
                  acc[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(
                      av[j], xr[4 * k4 + j][q], acc[q], 0, 0, 0);
                }
              }
            }
          }
#pragma unroll
          for (int g = 0; g < 4; ++g) {
            const int m0 = 8 * g + 4 * half;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
#pragma unroll
              for (int q = 0; q < S; ++q) {
                const float v = -acc[q][4 * g + e];
                const bool upd = v < bestv[q];
                bestv[q] = upd ? v : bestv[q];
                besti[q] = upd ? ((c0 + t) * 32 + m0 + e) : besti[q];
              }
            }
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const float ov = __shfl_xor(bestv[q], 32, kWave);
      const int oi = __shfl_xor(besti[q], 32, kWave);
      const bool take = (ov < bestv[q]) || (ov == bestv[q] && oi < besti[q]);
      bestv[q] = take ? ov : bestv[q];
      besti[q] = take ? oi : besti[q];
      if (half == 0 && valid[q]) labels[sidx[q]] = (I)besti[q];
    }
  }
}

// Centroid tiles in the FP32 kernel's LDS layout, so that a chunk is one
// contiguous copy: tile t holds C[2 ks + half][t*32 + col] at
// ((ks / 4) * 2 + half) * 128 + col * 4 + ks % 4, then -hn (-inf padded),
// the C operand the MFMA chain starts from.
template <int KS>
__global__ void __launch_bounds__(256)
    tiled_fp32_prep_krn(int d, int k, int n_ct, const float *__restrict__ C,
                        const float *__restrict__ hn, float *__restrict__ tiles) {
  constexpr int TF = tile_floats<KS>();
  const int64_t gid = blockIdx.x * 256ll + threadIdx.x;
  if (gid >= (int64_t)n_ct * TF) return;
  const int t = (int)(gid / TF), e = (int)(gid - (int64_t)t * TF);
  if (e < 2 * KS * 32) {
    // source order: feature-major (consecutive threads read consecutive
    // centroids of one feature row)
    const int f = e >> 5, cc = e & 31;
    const int c = t * 32 + cc;
    const int ks = f >> 1, hf = f & 1;
    tiles[(int64_t)t * TF + ((ks >> 2) * 2 + hf) * 128 + cc * 4 + (ks & 3)] =
        (f < d && c < k) ? C[(int64_t)f * k + c] : 0.f;
  } else {
    const int c = t * 32 + (e - 2 * KS * 32);
    tiles[(int64_t)t * TF + e] = (c < k) ? -hn[c] : -__builtin_huge_valf();
  }
}

// ---- candidate-tile re-score (BF16 path, at most 32 centroid tiles) ---------
// Each queued sample lists the centroid tiles that can hold its FP32 argmin;
// (sample, tile) pairs are bucketed by tile and each wave re-scores 32 pairs
// of one tile with the FP32 kernels' arithmetic, keeping the best (value,
// index) of a sample with a 64-bit atomicMin on an order-preserving key
// (ties: the lower index, as the FP32 path's strict < over ascending
// centroids).  Tiles outside a sample's list hold only centroids certified
// worse than its best, so the minimum over the listed tiles is the FP32
// path's label.

struct CandArgs {
  int n_segments, n_ct, d;
  int64_t seg_cap;
  const int *queue, *queue_counts;
  const unsigned *qmask;
  int *tile_cnt, *tile_cursor, *tile_start, *tile_witem;  // [n_ct + 1]
  int *pairs;
  int64_t pair_cap;
  int *overflow;                 // pairs did not fit: full re-score instead
  unsigned long long *keys;      // [n] per sample
  const float *Xr, *tiles;
  void *labels;
  int *n_rescored;
};

__device__ inline unsigned order_key(float v) {
  const unsigned b = __float_as_uint(v + 0.0f);  // -0 -> +0
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ void __launch_bounds__(256) cand_count_krn(CandArgs a) {
  __shared__ int cnt[32];
  const int seg = blockIdx.x;
  if (threadIdx.x < 32) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int m = a.queue_counts[seg];
  const int64_t base = (int64_t)seg * a.seg_cap;
  for (int j = threadIdx.x; j < m; j += 256) {
    unsigned bits = a.qmask[base + j];
    a.keys[a.queue[base + j]] = ~0ull;
    while (bits) {
      const int t = __ffs(bits) - 1;
      bits &= bits - 1u;
      atomicAdd(&cnt[t], 1);
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < a.n_ct && cnt[threadIdx.x] != 0)
    atomicAdd(&a.tile_cnt[threadIdx.x], cnt[threadIdx.x]);
}

__global__ void cand_scan_krn(CandArgs a) {
  if (threadIdx.x != 0) return;
  int run = 0, wrun = 0;
  for (int t = 0; t < a.n_ct; ++t) {
    const int c = a.tile_cnt[t];
    a.tile_start[t] = run;
    a.tile_witem[t] = wrun;
    run += c;
    wrun += (c + 31) / 32;
  }
  a.tile_start[a.n_ct] = run;
  a.tile_witem[a.n_ct] = wrun;
  a.overflow[0] = (int64_t)run > a.pair_cap ? 1 : 0;
  int q = 0;
  for (int g = 0; g < a.n_segments; ++g) q += a.queue_counts[g];
  a.n_rescored[0] = q;
}

__global__ void __launch_bounds__(256) cand_scatter_krn(CandArgs a) {
  __shared__ int cnt[32];
  if (a.overflow[0]) return;
  const int seg = blockIdx.x;
  if (threadIdx.x < 32) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int m = a.queue_counts[seg];
  const int64_t base = (int64_t)seg * a.seg_cap;
  for (int j = threadIdx.x; j < m; j += 256) {
    unsigned bits = a.qmask[base + j];
    while (bits) {
      const int t = __ffs(bits) - 1;
      bits &= bits - 1u;
      atomicAdd(&cnt[t], 1);
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < a.n_ct)
    cnt[threadIdx.x] = cnt[threadIdx.x] != 0
                           ? a.tile_start[threadIdx.x] +
                                 atomicAdd(&a.tile_cursor[threadIdx.x], cnt[threadIdx.x])
                           : 0;
  __syncthreads();
  for (int j = threadIdx.x; j < m; j += 256) {
    unsigned bits = a.qmask[base + j];
    const int smp = a.queue[base + j];
    while (bits) {
      const int t = __ffs(bits) - 1;
      bits &= bits - 1u;
      a.pairs[atomicAdd(&cnt[t], 1)] = smp;
    }
  }
}

template <int KS>
__global__ void __launch_bounds__(kTNT) cand_rescore_krn(CandArgs a) {
  constexpr int TF = tile_floats<KS>();
  if (a.overflow[0]) return;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int col = lane & 31, half = lane >> 5;
  const int total = a.tile_witem[a.n_ct];
  const int d = a.d;
  for (int wi = blockIdx.x * kTW + wave; wi < total; wi += gridDim.x * kTW) {
    int lo = 0, hi = a.n_ct - 1;  // the last tile whose first item is <= wi
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.tile_witem[mid] <= wi)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int t = lo;
    const int j0 = a.tile_start[t] + (wi - a.tile_witem[t]) * 32;
    const int cnt = min(32, a.tile_start[t + 1] - j0);
    const bool valid = col < cnt;
    const int smp = valid ? a.pairs[j0 + col] : 0;
    float xr[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int f = 2 * ks + half;
      xr[ks] = (valid && f < d) ? a.Xr[(int64_t)smp * d + f] : 0.f;
    }
    const float *At = a.tiles + (int64_t)t * TF;
    const float *Ht = At + 2 * KS * 32;
    f32x16 h;
#pragma unroll
    for (int gg = 0; gg < 4; ++gg) {
      const f32x4 h4 = *reinterpret_cast<const f32x4 *>(Ht + 8 * gg + 4 * half);
#pragma unroll
      for (int e = 0; e < 4; ++e) h[4 * gg + e] = h4[e];
    }
    f32x16 acc;
#pragma unroll
    for (int k4 = 0; k4 < KS / 4; ++k4) {
      const f32x4 av =
          *reinterpret_cast<const f32x4 *>(At + (k4 * 2 + half) * 128 + col * 4);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (k4 == 0 && j == 0) {
          acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[j], xr[0], h, 0, 0, 0);
        } else {
          acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[j], xr[4 * k4 + j], acc, 0,
                                                     0, 0);
        }
      }
    }
    float bestv = __builtin_huge_valf();
    int besti = 0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int m0 = 8 * g + 4 * half;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float v = -acc[4 * g + e];
        const bool upd = v < bestv;
        bestv = upd ? v : bestv;
        besti = upd ? (t * 32 + m0 + e) : besti;
      }
    }
    const float ov = __shfl_xor(bestv, 32, kWave);
    const int oi = __shfl_xor(besti, 32, kWave);
    const bool take = (ov < bestv) || (ov == bestv && oi < besti);
    bestv = take ? ov : bestv;
    besti = take ? oi : besti;
    if (half == 0 && valid && bestv < __builtin_huge_valf())
      atomicMin(&a.keys[smp],
                ((unsigned long long)order_key(bestv) << 32) | (unsigned)besti);
  }
}

template <typename I>
__global__ void __launch_bounds__(256) cand_finalize_krn(CandArgs a) {
  if (a.overflow[0]) return;
  const int seg = blockIdx.x;
  const int m = a.queue_counts[seg];
  const int64_t base = (int64_t)seg * a.seg_cap;
  I *labels = reinterpret_cast<I *>(a.labels);
  for (int j = threadIdx.x; j < m; j += 256) {
    const int smp = a.queue[base + j];
    const unsigned long long key = a.keys[smp];
    // no finite score anywhere: label 0, as the FP32 path
    labels[smp] = (key == ~0ull) ? (I)0 : (I)(unsigned)(key & 0xffffffffull);
  }
}

__global__ void __launch_bounds__(256) fill_zero_i32_krn(int m, int *p) {
  for (int i = threadIdx.x; i < m; i += 256) p[i] = 0;
}

// ---- BF16 assignment with certified margin ---------------------------------

inline int bf16_d_of(int64_t d) {
  if (d <= 32) return 32;
  if (d <= 64) return 64;
  if (d <= 96) return 96;
  return 128;
}
constexpr int kk_of(int D, int T) { return (T * D + 3 + 15) / 16; }
constexpr int bf16_s(int D, int T) { return D * T <= 128 ? 2 : 1; }
// LDS of the BF16 kernel: the per-tile maxima of every lane's sample blocks
// (bf16, for the candidate tiles), two centroid chunks of mixed_cpc tiles
// each (as large as kChunkBytes and the rest of LDS allow) and a counter
constexpr int mixed_tml_bytes(int D, int T) {
  return kTW * bf16_s(D, T) * 32 * kWave * (int)sizeof(unsigned short);
}
constexpr int mixed_cpc(int D, int T) {
  const int cap = (int)kLdsPerCu - mixed_tml_bytes(D, T) - 1024;
  const int chunk = kChunkBytes < cap / 2 ? kChunkBytes : cap / 2;
  return chunk / (kk_of(D, T) * 1024) < 1 ? 1 : chunk / (kk_of(D, T) * 1024);
}

struct TiledMixedArgs {
  int64_t n;
  int d, k, n_ct, idx_bits, fp32_len;
  const float *X;        // X_t (d, n)
  const bf16x8 *aop;     // [n_ct][KK][64] (+1 KB pad)
  const float *bound;    // [2]: max ||c||, max hn
  void *labels;
  int *queue;            // [block][seg_cap]
  int *queue_counts;     // [block]
  unsigned *qmask;       // [block][seg_cap]: candidate tiles (n_ct <= 32)
  int64_t seg_cap;
  float pinf;
  const int *skip;       // when non-null and *skip != 0 the launch does nothing
  // T == 1: the samples as BF16, sample-major (n, D), and their norm bounds
  // (prepared once per fit; required); the next round's operands are loaded
  // while the current round is scored
  const bf16x8 *xb;
  const float *xnb;
};

__device__ inline float max3f(float a, float b, float c) {
  float r;
  asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

template <int D, int T, typename I>
__global__ void __launch_bounds__(kTNT) __attribute__((amdgpu_waves_per_eu(2, 2)))
    tiled_mixed_krn(TiledMixedArgs a) {
  // an iteration enqueued ahead of a relocation: nothing to compute
  if (a.skip != nullptr && *a.skip != 0) return;
  constexpr int KD = D / 16, KK = (T * D + 3 + 15) / 16, XH = D / 2;
  constexpr int S = bf16_s(D, T);
  constexpr int CPC = mixed_cpc(D, T);
  constexpr int kBuf = CPC * KK * 64 * 16;  // bytes per buffer (1 KB multiple)
  extern __shared__ __align__(16) unsigned char lds_raw[];
  int &q_used = *reinterpret_cast<int *>(lds_raw + 2 * kBuf);
  unsigned short *tml = reinterpret_cast<unsigned short *>(lds_raw + 2 * kBuf + 16);
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1), wave = tid / kWave;
  const int col = lane & 31, half = lane >> 5;
  const int64_t n = a.n;
  const int d = a.d, n_ct = a.n_ct;
  I *labels = reinterpret_cast<I *>(a.labels);
  int *qseg = a.queue + (int64_t)blockIdx.x * a.seg_cap;
  if (tid == 0) q_used = 0;

  // a lane's row r (0..15) of a tile is packed into the 4 low mantissa
  // bits; the tile of the running best is tracked beside it
  constexpr unsigned kRowMask = 15u;
  const float cmax = a.bound[0], hnmax = a.bound[1];
  const bool finite_c = hnmax < 1e36f;
  const float pinf = a.pinf;
  constexpr float u = 1.0f / 256.0f;
  const float acc_err = (float)(T * D + 3 + a.fp32_len + 8) * 5.9604645e-8f;
  const float alpha = ((T == 1) ? (2.f * u + u * u) : (3.2f * u * u)) + acc_err;
  const float beta = acc_err;
  const float pack_err = ldexpf(1.0f, 4 - 23) * 1.01f;
  const bool vec = S == 2 && ((reinterpret_cast<uintptr_t>(a.X) & 7u) == 0) &&
                   (n % 2 == 0);

  constexpr int kSuper = 32 * S;
  const int64_t n_super = ceil_div<int64_t>(n, kSuper);
  const int64_t per_round = (int64_t)gridDim.x * kTW;
  const int64_t rounds = ceil_div<int64_t>(n_super, per_round);
  // A-operand chunks stream through two LDS buffers (chunk g + 1 copied
  // while chunk g is used; the sequence repeats every round)
  const int n_chunks = (n_ct + CPC - 1) / CPC;
  const int64_t total_chunks = rounds * n_chunks;
  const unsigned char *aop = reinterpret_cast<const unsigned char *>(a.aop);
  auto chunk_bytes = [&](int c) { return min(CPC, n_ct - c * CPC) * KK * 1024; };
  if (total_chunks > 0) lds_dma(aop, lds_raw, chunk_bytes(0), wave, lane);
  int64_t g = 0;
  // streamed BF16 operands (T == 1): lane (col, half) of sample s reads
  // features 16 kk + 8 half + 0..7 as one 16-byte load per k-step
  constexpr bool kXB = (T == 1);
  constexpr bool use_xb = kXB;
  bf16x8 nbh[kXB ? KD : 1][S];
  float nxn[S];
  auto load_xb = [&](int64_t rr) {
    const int64_t sbr = rr * per_round + (int64_t)blockIdx.x * kTW + wave;
    const int64_t sr = sbr * kSuper + (int64_t)S * col;
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const bool ok = sbr < n_super && sr + q < n;
      const bf16x8 *row = a.xb + (sr + q) * (D / 8) + half;
#pragma unroll
      for (int kk = 0; kk < (kXB ? KD : 1); ++kk) {
        if (ok) {
          nbh[kk][q] = row[2 * kk];
        } else {
#pragma unroll
          for (int j = 0; j < 8; ++j) nbh[kk][q][j] = (__bf16)0.f;
        }
      }
      nxn[q] = ok ? a.xnb[sr + q] : 0.f;
    }
  };
  if (use_xb && rounds > 0) load_xb(0);
  for (int64_t r = 0; r < rounds; ++r) {
    const int64_t sb = r * per_round + (int64_t)blockIdx.x * kTW + wave;
    const bool active = sb < n_super;
    const int64_t s0 = sb * kSuper + (int64_t)S * col;
    // lane features f(i) = 16 (i / 8) + 8 half + i % 8
    float xnorm[S];
    bf16x8 bhi[KD][S];
    bf16x8 blo[(T == 3) ? KD : 1][S];
    if constexpr (use_xb) {
#pragma unroll
      for (int q = 0; q < S; ++q) {
        xnorm[q] = nxn[q];
#pragma unroll
        for (int kk = 0; kk < KD; ++kk) bhi[kk][q] = nbh[kk < (kXB ? KD : 1) ? kk : 0][q];
      }
      // the next round's operands, in flight while this round is scored
      if (r + 1 < rounds) load_xb(r + 1);
    } else {
      float ss[S];
#pragma unroll
      for (int q = 0; q < S; ++q) ss[q] = 0.f;
      const bool whole = vec && active && (s0 + S <= n);
      // the row stride, opaque to loop-invariant code motion: otherwise the
      // XH row pointers are hoisted out of the round loop and stay live
      // (2 XH registers) through the MFMA loop
      int64_t n_row = n;
      asm volatile("" : "+s"(n_row));
      const float *xs = a.X + s0;
#pragma unroll
      for (int i = 0; i < XH; ++i) {
        const int f = 16 * (i / 8) + 8 * half + (i % 8);
        float xv[S];
        const float *row = xs + (int64_t)f * n_row;
        if (f < d && whole) {
          if constexpr (S == 2) {
            const float2 v = *reinterpret_cast<const float2 *>(row);
            xv[0] = v.x;
            xv[1] = v.y;
          }
        } else {
#pragma unroll
          for (int q = 0; q < S; ++q)
            xv[q] = (f < d && active && s0 + q < n) ? row[q] : 0.f;
        }
#pragma unroll
        for (int q = 0; q < S; ++q) {
          ss[q] = fmaf(xv[q], xv[q], ss[q]);
          const __bf16 h = (__bf16)xv[q];
          bhi[i / 8][q][i % 8] = h;
          if constexpr (T == 3) blo[i / 8][q][i % 8] = (__bf16)(xv[q] - (float)h);
        }
      }
#pragma unroll
      for (int q = 0; q < S; ++q) {
        const float t = ss[q] + __shfl_xor(ss[q], 32, kWave);
        xnorm[q] = sqrtf(t) * 1.0001f;
      }
    }
    bf16x8 bh;
#pragma unroll
    for (int j = 0; j < 8; ++j) bh[j] = (__bf16)((half == 0 && j < 3) ? -1.f : 0.f);

    float b1[S], b2[S];
    int bt[S];
    // candidate tiles of each sample (bit t: tile t), from the per-tile
    // maxima once the sample's best is known
    unsigned cmask[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
      b1[q] = -__builtin_huge_valf();
      b2[q] = -__builtin_huge_valf();
      bt[q] = 0;
      cmask[q] = 0u;
    }
    for (int ci = 0; ci < n_chunks; ++ci, ++g) {
      const int c0 = ci * CPC;
      const int cnt = min(CPC, n_ct - c0);
      wait_dma_and_sync();
      if (g + 1 < total_chunks) {
        const int cn = (int)((g + 1) % n_chunks);
        lds_dma(aop + (size_t)cn * CPC * KK * 1024, lds_raw + ((g + 1) & 1) * kBuf,
                chunk_bytes(cn), wave, lane);
      }
      const bf16x8 *Al = reinterpret_cast<const bf16x8 *>(lds_raw + (g & 1) * kBuf);
      if (active) {
        for (int t = 0; t < cnt; ++t) {
          const bf16x8 *At = Al + t * KK * 64 + lane;
          f32x16 acc[S];
          // The tile's A operands are read from LDS in groups of up to kG,
          // each group issued back to back and waited for once: an empty
          // statement that takes the group's values keeps the scheduler from
          // moving each read next to its MFMA (which exposes the LDS latency
          // once per k-step).
          constexpr int kG = KK < 10 ? KK : 9;
#pragma unroll
          for (int k0 = 0; k0 < KK; k0 += kG) {
            bf16x8 ava[kG];
#pragma unroll
            for (int j = 0; j < kG; ++j)
              if (k0 + j < KK) ava[j] = At[(k0 + j) * 64];
#pragma unroll
            for (int j = 0; j < kG; ++j)
              if (k0 + j < KK) asm volatile("" : : "v"(ava[j]));
#pragma unroll
            for (int j = 0; j < kG; ++j) {
              const int kk = k0 + j;
              if (kk >= KK) break;
              const bf16x8 av = ava[j];
#pragma unroll
              for (int q = 0; q < S; ++q) {
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
          }
          const int ct = c0 + t;
          float tmax[S], b1_in[S];
#pragma unroll
          for (int q = 0; q < S; ++q) {
            tmax[q] = -__builtin_huge_valf();
            b1_in[q] = b1[q];
          }
#pragma unroll
          for (int rr = 0; rr < 16; rr += 2) {
#pragma unroll
            for (int q = 0; q < S; ++q) {
              const float uu = __uint_as_float(
                  (__float_as_uint(acc[q][rr]) & ~kRowMask) | (unsigned)rr);
              const float vv = __uint_as_float(
                  (__float_as_uint(acc[q][rr + 1]) & ~kRowMask) | (unsigned)(rr + 1));
              const float m = __builtin_amdgcn_fmed3f(b1[q], uu, vv);
              b1[q] = max3f(b1[q], uu, vv);
              b2[q] = __builtin_amdgcn_fmed3f(b2[q], m, pinf);
              tmax[q] = max3f(tmax[q], uu, vv);
            }
          }
#pragma unroll
          for (int q = 0; q < S; ++q) bt[q] = (b1[q] != b1_in[q]) ? ct : bt[q];
          // the tile's best, rounded toward +inf to bf16 (never below it)
#pragma unroll
          for (int q = 0; q < S; ++q)
            if (a.qmask != nullptr && ct < 32) {
              const unsigned b = __float_as_uint(tmax[q]);
              const unsigned up = (b & 0x80000000u) ? (b >> 16) : ((b + 0xffffu) >> 16);
              tml[((wave * S + q) * 32 + ct) * kWave + lane] = (unsigned short)up;
            }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const unsigned rb = __float_as_uint(b1[q]) & kRowMask;
      unsigned lab = (unsigned)(bt[q] * 32) + 8u * (rb >> 2) + 4u * half + (rb & 3u);
      const float o1 = __shfl_xor(b1[q], 32, kWave);
      const float o2 = __shfl_xor(b2[q], 32, kWave);
      const unsigned ol = (unsigned)__shfl_xor((int)lab, 32, kWave);
      const float n2 = fmaxf(fminf(b1[q], o1), fmaxf(b2[q], o2));
      lab = (o1 > b1[q]) ? ol : lab;
      b1[q] = fmaxf(b1[q], o1);
      b2[q] = n2;
      const float e = xnorm[q] * cmax * alpha + hnmax * beta +
                      pack_err * (fabsf(b1[q]) + fabsf(b2[q]));
      const bool amb = !(b1[q] - b2[q] > 2.f * e) || !(xnorm[q] * cmax < 1e36f) ||
                       !finite_c || lab >= (unsigned)a.k;
      const int64_t s = s0 + q;
      const bool valid = active && s < n;
      if (half == 0 && valid && !amb) labels[s] = (I)lab;
      const bool enq = (half == 0) && valid && amb;
      {
        // candidate tiles against the final best: a centroid can be the FP32
        // argmin only if its score is within 2 (E1 + E2) of the best; with
        // the packing error of both values this is 2 e_part + 6 pack M
        if (a.qmask != nullptr) {
          const float big = 1.01f * (xnorm[q] * cmax + hnmax);
          const float thr = b1[q] - (2.f * (xnorm[q] * cmax * alpha + hnmax * beta) +
                                     6.f * pack_err * big);
          unsigned bits = 0u;
          const int nt = min(n_ct, 32);
          for (int t = 0; t < nt; ++t) {
            const float v = __uint_as_float(
                (unsigned)tml[((wave * S + q) * 32 + t) * kWave + lane] << 16);
            if (v >= thr) bits |= 1u << t;
          }
          cmask[q] = bits;
        }
      }
      const unsigned tiles_any = cmask[q] | (unsigned)__shfl_xor((int)cmask[q], 32, kWave);
      const uint64_t mask = __ballot(enq);
      if (mask != 0ull) {
        const int c = __popcll(mask);
        const int leader = __ffsll((long long)mask) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(&q_used, c);
        base = __shfl(base, leader, kWave);
        if (enq) {
          const int rank = __popcll(mask & ((1ull << lane) - 1ull));
          qseg[base + rank] = (int)s;
          if (a.qmask) a.qmask[(int64_t)blockIdx.x * a.seg_cap + base + rank] = tiles_any;
        }
      }
    }
  }
  __syncthreads();
  if (tid == 0) a.queue_counts[blockIdx.x] = q_used;
}

// ---- FP8 (e4m3) assignment with per-tile scales and certified margin ----------
//
// The cross term on the matrix core in FP8 (v_mfma_f32_32x32x16_fp8_fp8, FP32
// accumulation): centroid tile t enters as e4m3(c s_t), its own power-of-two
// scale s_t keeping |c s_t| <= kF8Target, the samples as e4m3(x s_x) with one
// power-of-two scale for the fit.  The accumulator starts from
// -hn s_t s_x (FP32) and the result is unscaled exactly by 1 / (s_t s_x), so
// D = sum_f x~_f c~_f - hn, argmax D = argmin (hn - c.x).  The decision and
// the FP32 re-score of undecided samples are the BF16 kernel's; the error
// bound is e4m3's: per operand |e4m3(v) - v| <= u |v| + eta (u = 2^-4 for
// round-to-nearest with 3 mantissa bits; eta = 2^-10, half the smallest
// subnormal spacing), so
//   |D~ - D| <= ||x|| cmax (2u + u^2) + sqrt(D) (eta_x cmax + eta_c ||x||)(1 + u)
//               + D eta_x eta_c + FP32 accumulation,
// with eta_x = eta / s_x and eta_c = eta / min_t s_t in data units.
constexpr int kF8 = 8;                    // the "terms" value of the FP8 kernels
constexpr float kF8Target = 224.f;        // scaled magnitudes (both encodings' range)
constexpr float kF8U = 0.0625f;           // e4m3 unit roundoff
constexpr float kF8Eta = 1.0f / 1024.f;   // half the smallest subnormal spacing
typedef long long2v __attribute__((ext_vector_type(2)));

constexpr int f8_slots(int D) { return D / 32 + 1; }   // 1 KB A-operand slots per tile
constexpr int f8_s() { return 2; }                     // sample blocks per wave
constexpr int f8_tml_bytes() {
  return kTW * f8_s() * 32 * kWave * (int)sizeof(unsigned short);
}
constexpr int f8_cpc(int D) {
  const int cap = (int)kLdsPerCu - f8_tml_bytes() - 1024;
  const int chunk = kChunkBytes < cap / 2 ? kChunkBytes : cap / 2;
  return chunk / (f8_slots(D) * 1024) < 1 ? 1 : chunk / (f8_slots(D) * 1024);
}

// Largest power of two p with m p <= kF8Target (1 for m = 0 or non-finite m).
__device__ inline float f8_scale(float m) {
  if (!(m > 0.f) || !(m < 1e30f)) return 1.f;
  int e;
  frexpf(kF8Target / m, &e);   // kF8Target / m in [2^(e-1), 2^e)
  return ldexpf(1.f, e - 1);
}

// e4m3 of eight values, element j in byte j
__device__ inline long f8_pack8(const float (&v)[8]) {
  int lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
  lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], lo, true);
  int hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[4], v[5], 0, false);
  hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[6], v[7], hi, true);
  return (long)(unsigned)lo | ((long)(unsigned)hi << 32);
}

template <int D, typename I>
__global__ void __launch_bounds__(kTNT) __attribute__((amdgpu_waves_per_eu(2, 2)))
    tiled_fp8_krn(TiledMixedArgs a) {
  if (a.skip != nullptr && *a.skip != 0) return;
  constexpr int KD = D / 16, KH = D / 32, KK = f8_slots(D);
  constexpr int S = f8_s();
  constexpr int CPC = f8_cpc(D);
  constexpr int kBuf = CPC * KK * 1024;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  int &q_used = *reinterpret_cast<int *>(lds_raw + 2 * kBuf);
  unsigned short *tml = reinterpret_cast<unsigned short *>(lds_raw + 2 * kBuf + 16);
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1), wave = tid / kWave;
  const int col = lane & 31, half = lane >> 5;
  const int64_t n = a.n;
  const int n_ct = a.n_ct;
  I *labels = reinterpret_cast<I *>(a.labels);
  int *qseg = a.queue + (int64_t)blockIdx.x * a.seg_cap;
  if (tid == 0) q_used = 0;

  // a lane's row r (0..15) of a tile is packed into the 4 low mantissa
  // bits; the tile of the running best is tracked beside it
  constexpr unsigned kRowMask = 15u;
  // bound: [0] max ||c||, [1] max hn, [2] eta / min s_t, [3] s_x, [4] eta / s_x,
  // [5] max s_t
  const float cmax = a.bound[0], hnmax = a.bound[1];
  const float eta_c = a.bound[2], sx = a.bound[3], eta_x = a.bound[4];
  const float smax = a.bound[5];
  const bool finite_c = hnmax < 1e30f && (cmax + hnmax) * smax * sx < 1e30f;
  const float pinf = a.pinf;
  const float acc_err = (float)(D + 2 + a.fp32_len + 8) * 5.9604645e-8f;
  const float alpha = 2.f * kF8U + kF8U * kF8U + acc_err * (1.f + kF8U) * (1.f + kF8U);
  const float beta = acc_err;
  const float rootd = sqrtf((float)D);
  const float pack_err = ldexpf(1.0f, 4 - 23) * 1.01f;

  constexpr int kSuper = 32 * S;
  const int64_t n_super = ceil_div<int64_t>(n, kSuper);
  const int64_t per_round = (int64_t)gridDim.x * kTW;
  const int64_t rounds = ceil_div<int64_t>(n_super, per_round);
  const int n_chunks = (n_ct + CPC - 1) / CPC;
  const int64_t total_chunks = rounds * n_chunks;
  const unsigned char *aop = reinterpret_cast<const unsigned char *>(a.aop);
  auto chunk_bytes = [&](int c) { return min(CPC, n_ct - c * CPC) * KK * 1024; };
  if (total_chunks > 0) lds_dma(aop, lds_raw, chunk_bytes(0), wave, lane);
  int64_t g = 0;
  // streamed FP8 samples: lane (col, half) of sample s reads the D / 2 bytes
  // of its half (k-steps 2 j and 2 j + 1 in one 16-byte load)
  const unsigned char *x8 = reinterpret_cast<const unsigned char *>(a.xb);
  long2v nb[KH][S];
  float nxn[S];
  auto load_x8 = [&](int64_t rr) {
    const int64_t sbr = rr * per_round + (int64_t)blockIdx.x * kTW + wave;
    const int64_t sr = sbr * kSuper + (int64_t)S * col;
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const bool ok = sbr < n_super && sr + q < n;
      const long2v *row = reinterpret_cast<const long2v *>(x8 + (sr + q) * D + half * (D / 2));
#pragma unroll
      for (int j = 0; j < KH; ++j) {
        if (ok) {
          nb[j][q] = row[j];
        } else {
          nb[j][q][0] = 0;
          nb[j][q][1] = 0;
        }
      }
      nxn[q] = ok ? a.xnb[sr + q] : 0.f;
    }
  };
  if (rounds > 0) load_x8(0);
  for (int64_t r = 0; r < rounds; ++r) {
    const int64_t sb = r * per_round + (int64_t)blockIdx.x * kTW + wave;
    const bool active = sb < n_super;
    const int64_t s0 = sb * kSuper + (int64_t)S * col;
    float xnorm[S];
    long2v bx[KH][S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
      xnorm[q] = nxn[q];
#pragma unroll
      for (int j = 0; j < KH; ++j) bx[j][q] = nb[j][q];
    }
    if (r + 1 < rounds) load_x8(r + 1);
    float b1[S], b2[S];
    int bt[S];
    unsigned cmask[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
      b1[q] = -__builtin_huge_valf();
      b2[q] = -__builtin_huge_valf();
      bt[q] = 0;
      cmask[q] = 0u;
    }
    for (int ci = 0; ci < n_chunks; ++ci, ++g) {
      const int c0 = ci * CPC;
      const int cnt = min(CPC, n_ct - c0);
      wait_dma_and_sync();
      if (g + 1 < total_chunks) {
        const int cn = (int)((g + 1) % n_chunks);
        lds_dma(aop + (size_t)cn * CPC * KK * 1024, lds_raw + ((g + 1) & 1) * kBuf,
                chunk_bytes(cn), wave, lane);
      }
      const unsigned char *Al = lds_raw + (g & 1) * kBuf;
      if (active) {
        for (int t = 0; t < cnt; ++t) {
          const unsigned char *At = Al + (size_t)t * KK * 1024;
          const float *Ht = reinterpret_cast<const float *>(At + KH * 1024);
          long2v av[KH];
#pragma unroll
          for (int j = 0; j < KH; ++j)
            av[j] = *reinterpret_cast<const long2v *>(At + j * 1024 + lane * 16);
          f32x16 init;
#pragma unroll
          for (int gg = 0; gg < 4; ++gg) {
            const f32x4 h4 = *reinterpret_cast<const f32x4 *>(Ht + 8 * gg + 4 * half);
#pragma unroll
            for (int e = 0; e < 4; ++e) init[4 * gg + e] = h4[e];
          }
          const float inv = Ht[32];
          f32x16 acc[S];
#pragma unroll
          for (int q = 0; q < S; ++q) {
            acc[q] = init;
#pragma unroll
            for (int kk = 0; kk < KD; ++kk)
              acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(
                  av[kk / 2][kk % 2], bx[kk / 2][q][kk % 2], acc[q], 0, 0, 0);
          }
          const int ct = c0 + t;
          float tmax[S], b1_in[S];
#pragma unroll
          for (int q = 0; q < S; ++q) {
            tmax[q] = -__builtin_huge_valf();
            b1_in[q] = b1[q];
          }
#pragma unroll
          for (int rr = 0; rr < 16; rr += 2) {
#pragma unroll
            for (int q = 0; q < S; ++q) {
              const float uu = __uint_as_float(
                  (__float_as_uint(acc[q][rr] * inv) & ~kRowMask) | (unsigned)rr);
              const float vv = __uint_as_float(
                  (__float_as_uint(acc[q][rr + 1] * inv) & ~kRowMask) | (unsigned)(rr + 1));
              const float m = __builtin_amdgcn_fmed3f(b1[q], uu, vv);
              b1[q] = max3f(b1[q], uu, vv);
              b2[q] = __builtin_amdgcn_fmed3f(b2[q], m, pinf);
              tmax[q] = max3f(tmax[q], uu, vv);
            }
          }
#pragma unroll
          for (int q = 0; q < S; ++q) bt[q] = (b1[q] != b1_in[q]) ? ct : bt[q];
#pragma unroll
          for (int q = 0; q < S; ++q)
            if (a.qmask != nullptr && ct < 32) {
              const unsigned b = __float_as_uint(tmax[q]);
              const unsigned up = (b & 0x80000000u) ? (b >> 16) : ((b + 0xffffu) >> 16);
              tml[((wave * S + q) * 32 + ct) * kWave + lane] = (unsigned short)up;
            }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const unsigned rb = __float_as_uint(b1[q]) & kRowMask;
      unsigned lab = (unsigned)(bt[q] * 32) + 8u * (rb >> 2) + 4u * half + (rb & 3u);
      const float o1 = __shfl_xor(b1[q], 32, kWave);
      const float o2 = __shfl_xor(b2[q], 32, kWave);
      const unsigned ol = (unsigned)__shfl_xor((int)lab, 32, kWave);
      const float n2 = fmaxf(fminf(b1[q], o1), fmaxf(b2[q], o2));
      lab = (o1 > b1[q]) ? ol : lab;
      b1[q] = fmaxf(b1[q], o1);
      b2[q] = n2;
      const float e_part = xnorm[q] * cmax * alpha + hnmax * beta +
                           rootd * (eta_x * cmax + eta_c * xnorm[q]) * (1.f + kF8U) +
                           (float)D * eta_x * eta_c;
      const float e = e_part + pack_err * (fabsf(b1[q]) + fabsf(b2[q]));
      const bool amb = !(b1[q] - b2[q] > 2.f * e) || !(xnorm[q] * cmax < 1e30f) ||
                       !finite_c || lab >= (unsigned)a.k;
      const int64_t s = s0 + q;
      const bool valid = active && s < n;
      if (half == 0 && valid && !amb) labels[s] = (I)lab;
      const bool enq = (half == 0) && valid && amb;
      if (a.qmask != nullptr) {
        const float big = 1.01f * (xnorm[q] * cmax + hnmax);
        const float thr = b1[q] - (2.f * e_part + 6.f * pack_err * big);
        unsigned bits = 0u;
        const int nt = min(n_ct, 32);
        for (int t = 0; t < nt; ++t) {
          const float v = __uint_as_float(
              (unsigned)tml[((wave * S + q) * 32 + t) * kWave + lane] << 16);
          if (v >= thr) bits |= 1u << t;
        }
        cmask[q] = bits;
      }
      const unsigned tiles_any = cmask[q] | (unsigned)__shfl_xor((int)cmask[q], 32, kWave);
      const uint64_t mask = __ballot(enq);
      if (mask != 0ull) {
        const int c = __popcll(mask);
        const int leader = __ffsll((long long)mask) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(&q_used, c);
        base = __shfl(base, leader, kWave);
        if (enq) {
          const int rank = __popcll(mask & ((1ull << lane) - 1ull));
          qseg[base + rank] = (int)s;
          if (a.qmask) a.qmask[(int64_t)blockIdx.x * a.seg_cap + base + rank] = tiles_any;
        }
      }
    }
  }
  __syncthreads();
  if (tid == 0) a.queue_counts[blockIdx.x] = q_used;
}

// FP8 A operands of tile blockIdx.x (one block per tile): slot j < D / 32 holds,
// for lane l (centroid ct 32 + l % 32), e4m3(c s_t) of k-steps 2 j and 2 j + 1
// (K positions 16 kk + 8 (l / 32) + 0..7 in bytes 8 (kk % 2) + 0..7); slot
// D / 32 holds the chain's start -hn s_t s_x for the 32 rows, then 1 / (s_t s_x).
// bound[2] and bound[5] gather eta / s_t and s_t over the tiles (the launcher
// clears them); bound[3] = s_x.
__global__ void __launch_bounds__(256)
    tiled_fp8_prep_krn(int d, int k, int D, const float *__restrict__ C,
                       const float *__restrict__ hn, unsigned char *__restrict__ aop,
                       float *bound) {
  __shared__ float red[256 / kWave];
  __shared__ float s_tile;
  const int ct = blockIdx.x, tid = threadIdx.x;
  const int KH = D / 32, KK = KH + 1;
  float m = 0.f;
  for (int i = tid; i < 32 * d; i += 256) {
    const int f = i >> 5, c = ct * 32 + (i & 31);
    if (c < k) m = fmaxf(m, fabsf(C[(int64_t)f * k + c]));
  }
  for (int o = kWave / 2; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, kWave));
  if ((tid & (kWave - 1)) == 0) red[tid / kWave] = m;
  __syncthreads();
  if (tid == 0) {
    float t = 0.f;
    for (int i = 0; i < 256 / kWave; ++i) t = fmaxf(t, red[i]);
    s_tile = f8_scale(t);
    atomicMax(reinterpret_cast<unsigned *>(&bound[2]), __float_as_uint(kF8Eta / s_tile));
    atomicMax(reinterpret_cast<unsigned *>(&bound[5]), __float_as_uint(s_tile));
  }
  __syncthreads();
  const float st = s_tile, sx = bound[3];
  unsigned char *tile = aop + (size_t)ct * KK * 1024;
  for (int gi = tid; gi < KH * 64; gi += 256) {
    const int j = gi >> 6, l = gi & 63;
    const int c = ct * 32 + (l & 31), hf = l >> 5;
    long2v out;
#pragma unroll
    for (int h2 = 0; h2 < 2; ++h2) {
      const int kk = 2 * j + h2;
      float v[8];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const int f = 16 * kk + 8 * hf + jj;
        v[jj] = (f < d && c < k) ? C[(int64_t)f * k + c] * st : 0.f;
      }
      out[h2] = f8_pack8(v);
    }
    *reinterpret_cast<long2v *>(tile + j * 1024 + l * 16) = out;
  }
  float *hs = reinterpret_cast<float *>(tile + KH * 1024);
  const float sc = st * sx;
  for (int i = tid; i < 256; i += 256) {
    float v = 0.f;
    if (i < 32) {
      const int c = ct * 32 + i;
      // padding rows: far below every real score once unscaled (>= -3e38)
      v = c < k ? -hn[c] * sc : fmaxf(-3.0e38f, -1.0e30f * sc);
    } else if (i == 32) {
      v = 1.f / sc;
    }
    hs[i] = v;
  }
}

// max |x| over X_t as a bit pattern (non-negative floats order like their bits)
__global__ void __launch_bounds__(256)
    tiled_x8_max_krn(int64_t count, const float *__restrict__ X_t, unsigned *maxbits) {
  unsigned m = 0u;
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * 256)
    m = max(m, __float_as_uint(X_t[i]) & 0x7fffffffu);
  for (int o = kWave / 2; o > 0; o >>= 1)
    m = max(m, (unsigned)__shfl_xor((int)m, o, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0) atomicMax(maxbits, m);
}

// x8[s] = e4m3(x_s s_x) laid out per half of the lane pair (bytes
// h D/2 + 8 kk + j = feature 16 kk + 8 h + j), xnb[s] = 1.0001 ||x_s||;
// block 0 writes bound[3] = s_x and bound[4] = eta / s_x.
__global__ void __launch_bounds__(256)
    tiled_x8_prep_krn(int64_t n, int d, int D, const float *__restrict__ X_t,
                      const unsigned *maxbits, unsigned char *__restrict__ x8,
                      float *__restrict__ xnb, float *bound) {
  const float sx = f8_scale(__uint_as_float(*maxbits));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    bound[3] = sx;
    bound[4] = kF8Eta / sx;
  }
  const int64_t s = blockIdx.x * 256ll + threadIdx.x;
  if (s >= n) return;
  float ss = 0.f;
  for (int hf = 0; hf < 2; ++hf)
    for (int kk = 0; kk < D / 16; ++kk) {
      float v[8];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const int f = 16 * kk + 8 * hf + jj;
        const float x = (f < d) ? X_t[(int64_t)f * n + s] : 0.f;
        ss = fmaf(x, x, ss);
        v[jj] = x * sx;
      }
      *reinterpret_cast<long *>(x8 + s * D + hf * (D / 2) + 8 * kk) = f8_pack8(v);
    }
  xnb[s] = sqrtf(ss) * 1.0001f;
}

// e4m3 encode + decode of in[i] (the FP8 kernels' conversion), for tests of
// the rounding the error bound assumes
__global__ void __launch_bounds__(256)
    f8_roundtrip_krn(int64_t n, const float *__restrict__ in, float *__restrict__ out) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  const int w = __builtin_amdgcn_cvt_pk_fp8_f32(in[i], 0.f, 0, false);
  out[i] = __builtin_amdgcn_cvt_f32_fp8(w, 0);
}

// A operands of the BF16 kernel for a runtime bucket D: element j of lane l
// for (tile ct, k-step kk) is centroid ct*32 + l%32 at K position
// p = 16 kk + 8 (l / 32) + j: p < T*D the split centroid features, then the
// three bf16 terms of hn (padding centroids: 3e38, far below every real
// score that passes the kernel's finiteness guards, so they never win).
__global__ void __launch_bounds__(256)
    tiled_prep_krn(int d, int k, int D, int T, int KK, int n_ct,
                   const float *__restrict__ C, const float *__restrict__ hn,
                   bf16x8 *__restrict__ aop) {
  const int64_t total = (int64_t)n_ct * KK * 64;
  const int64_t gid = blockIdx.x * 256ll + threadIdx.x;
  if (gid >= total) return;
  const int lane = (int)(gid % 64);
  const int kk = (int)((gid / 64) % KK);
  const int ct = (int)(gid / (64 * KK));
  const int c = ct * 32 + lane % 32;
  bf16x8 v;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int p = 16 * kk + 8 * (lane / 32) + j;
    float val = 0.f;
    if (p < T * D) {
      const int seg = p / D, f = p % D;
      const float cv = (f < d && c < k) ? C[(int64_t)f * k + c] : 0.f;
      const float hi = (float)(__bf16)cv;
      val = (T == 3 && seg == 1) ? (float)(__bf16)(cv - hi) : hi;
    } else if (p < T * D + 3) {
      const int t = p - T * D;
      if (c < k) {
        const float h = hn[c];
        const float h1 = (float)(__bf16)h;
        const float r1 = h - h1;
        const float h2 = (float)(__bf16)r1;
        const float h3 = (float)(__bf16)(r1 - h2);
        val = (t == 0) ? h1 : (t == 1 ? h2 : h3);
      } else {
        val = (t == 0) ? 3.0e38f : 0.f;
      }
    }
    v[j] = (__bf16)val;
  }
  aop[gid] = v;
}

// xb[s] = the sample's features as BF16 (row of D, zero past d) and xnb[s]
// = 1.0001 ||x_s|| (FP32; an upper bound of the norm), for the BF16 kernel's
// streamed operands.  One thread per sample; the features are read
// feature-major (coalesced over samples) and written 16 bytes at a time.
__global__ void __launch_bounds__(256)
    tiled_xb_prep_krn(int64_t n, int d, int D, const float *__restrict__ X_t,
                      bf16x8 *__restrict__ xb, float *__restrict__ xnb) {
  const int64_t s = blockIdx.x * 256ll + threadIdx.x;
  if (s >= n) return;
  float ss = 0.f;
  for (int f0 = 0; f0 < D; f0 += 8) {
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int f = f0 + j;
      const float x = (f < d) ? X_t[(int64_t)f * n + s] : 0.f;
      ss = fmaf(x, x, ss);
      v[j] = (__bf16)x;
    }
    xb[s * (D / 8) + f0 / 8] = v;
  }
  xnb[s] = sqrtf(ss) * 1.0001f;
}

__global__ void __launch_bounds__(256)
    tiled_bound_krn(int k, const float *__restrict__ hn, float *bound) {
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

// ---- transpose and update ----------------------------------------------------

__global__ void __launch_bounds__(256)
    transpose_krn(int64_t n, int d, const float *__restrict__ X_t,
                  float *__restrict__ X_r) {
  __shared__ float tile[32][33];
  const int64_t s0 = (int64_t)blockIdx.x * 32;
  const int f0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 8 rows per pass
  for (int r = ty; r < 32; r += 8) {
    const int f = f0 + r;
    const int64_t s = s0 + tx;
    tile[r][tx] = (f < d && s < n) ? X_t[(int64_t)f * n + s] : 0.f;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int64_t s = s0 + r;
    const int f = f0 + tx;
    if (s < n && f < d) X_r[s * d + f] = tile[tx][r];
  }
}

// totals[c] += number of samples with label c in block b's chunk (one
// global atomic per block and non-empty cluster)
template <typename I>
__global__ void __launch_bounds__(256)
    label_hist_krn(int64_t n, int k, int64_t chunk, const I *__restrict__ labels,
                   int *__restrict__ totals, const int *skip) {
  if (skip != nullptr && *skip != 0) return;
  extern __shared__ int h[];
  for (int c = threadIdx.x; c < k; c += 256) h[c] = 0;
  __syncthreads();
  const int64_t beg = (int64_t)blockIdx.x * chunk;
  const int64_t end = min(n, beg + chunk);
  for (int64_t s = beg + threadIdx.x; s < end; s += 256)
    atomicAdd(&h[(int)labels[s]], 1);
  __syncthreads();
  for (int c = threadIdx.x; c < k; c += 256)
    if (h[c] != 0) atomicAdd(&totals[c], h[c]);
}

// start[c] = exclusive scan of totals (start[k] = n); items[c] = exclusive
// scan of ceil(totals[c] / members) (items[k] = number of work items).  One
// block.
__global__ void __launch_bounds__(1024)
    cluster_scan_krn(int k, int members, const int *__restrict__ totals,
                     int *__restrict__ start, int *__restrict__ items,
                     const int *skip) {
  if (skip != nullptr && *skip != 0) return;
  __shared__ int wsum_a[1024 / kWave], wsum_b[1024 / kWave];
  __shared__ int carry_a, carry_b;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  if (tid == 0) {
    carry_a = 0;
    carry_b = 0;
  }
  __syncthreads();
  for (int c0 = 0; c0 < k; c0 += 1024) {
    const int c = c0 + tid;
    const int va = (c < k) ? totals[c] : 0;
    const int vb = (c < k) ? (va + members - 1) / members : 0;
    int ia = va, ib = vb;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const int ta = __shfl_up(ia, o, kWave), tb = __shfl_up(ib, o, kWave);
      if (lane >= o) {
        ia += ta;
        ib += tb;
      }
    }
    if (lane == kWave - 1) {
      wsum_a[wave] = ia;
      wsum_b[wave] = ib;
    }
    __syncthreads();
    int pa = 0, pb = 0;
    for (int w2 = 0; w2 < wave; ++w2) {
      pa += wsum_a[w2];
      pb += wsum_b[w2];
    }
    if (c < k) {
      start[c] = carry_a + pa + ia - va;
      items[c] = carry_b + pb + ib - vb;
    }
    __syncthreads();
    if (tid == 1023) {
      carry_a += pa + ia;
      carry_b += pb + ib;
    }
    __syncthreads();
  }
  if (tid == 0) {
    start[k] = carry_a;
    items[k] = carry_b;
  }
}

// perm[start[c] + base + rank] = s for the samples of block b's chunk with
// label c, where base is the block's share of cluster c claimed from
// cursor[c] (the order inside a cluster is irrelevant: the sums are exact)
template <typename I>
__global__ void __launch_bounds__(256)
    scatter_krn(int64_t n, int k, int64_t chunk, const I *__restrict__ labels,
                const int *__restrict__ start, int *__restrict__ cursor,
                int *__restrict__ perm, const int *skip) {
  if (skip != nullptr && *skip != 0) return;
  extern __shared__ int h[];
  for (int c = threadIdx.x; c < k; c += 256) h[c] = 0;
  __syncthreads();
  const int64_t beg = (int64_t)blockIdx.x * chunk;
  const int64_t end = min(n, beg + chunk);
  for (int64_t s = beg + threadIdx.x; s < end; s += 256)
    atomicAdd(&h[(int)labels[s]], 1);
  __syncthreads();
  for (int c = threadIdx.x; c < k; c += 256)
    h[c] = (h[c] != 0) ? start[c] + atomicAdd(&cursor[c], h[c]) : 0;
  __syncthreads();
  for (int64_t s = beg + threadIdx.x; s < end; s += 256) {
    const int pos = atomicAdd(&h[(int)labels[s]], 1);
    perm[pos] = (int)s;
  }
}

__global__ void __launch_bounds__(256) fill_i32_krn(int64_t m, int *p) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * 256)
    p[i] = 0;
}

__global__ void __launch_bounds__(256) fill_u64_krn(int64_t m, u64 *p) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * 256)
    p[i] = 0ull;
}

// One work item = up to kMembers members of one cluster; thread f sums
// rint((x_f w) 2^E_f) (f < d) or rint(w 2^E_w) (f == d) over them and adds
// the int64 total into sums[f * k + c].
__global__ void __launch_bounds__(320)
    segment_sum_krn(int d, int k, const float *__restrict__ X_r,
                    const float *__restrict__ w, const float *__restrict__ scale,
                    const int *__restrict__ perm, const int *__restrict__ start,
                    const int *__restrict__ items, u64 *__restrict__ sums,
                    const int *skip) {
  __shared__ int sidx[kMembers];
  __shared__ float sw[kMembers];
  if (skip != nullptr && *skip != 0) return;
  const int item = blockIdx.x;
  if (item >= items[k]) return;
  int lo = 0, hi = k - 1;  // the last cluster whose first item is <= item
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid] <= item)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int c = lo;
  const int m0 = start[c] + (item - items[c]) * kMembers;
  const int cnt = min(kMembers, start[c + 1] - m0);
  for (int j = threadIdx.x; j < cnt; j += blockDim.x) {
    const int s = perm[m0 + j];
    sidx[j] = s;
    sw[j] = w[s];
  }
  __syncthreads();
  for (int f = threadIdx.x; f <= d; f += blockDim.x) {
    const float sc = scale[f];
    long long acc = 0;
    if (f < d) {
      constexpr int kU = 16;  // row gathers in flight per thread (integer sums: any order)
      int j = 0;
      for (; j + kU <= cnt; j += kU) {
        float xv[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) xv[u] = X_r[(int64_t)sidx[j + u] * d + f];
#pragma unroll
        for (int u = 0; u < kU; ++u) acc += __float2int_rn((xv[u] * sw[j + u]) * sc);
      }
      for (; j < cnt; ++j)
        acc += __float2int_rn((X_r[(int64_t)sidx[j] * d + f] * sw[j]) * sc);
    } else {
      for (int j = 0; j < cnt; ++j) acc += __float2int_rn(sw[j] * sc);
    }
    if (acc != 0) atomicAdd(&sums[(int64_t)f * k + c], (u64)acc);
  }
}

// ---- plans --------------------------------------------------------------------

using TiledFn = void (*)(TiledArgs);
using TiledMixedFn = void (*)(TiledMixedArgs);

template <typename I, bool GATHER>
TiledFn pick_fp32(int64_t d) {
  switch (ks_of(d)) {
    case 16: return &tiled_fp32_krn<16, I, GATHER>;
    case 32: return &tiled_fp32_krn<32, I, GATHER>;
    default: return &tiled_fp32_krn<64, I, GATHER>;
  }
}

}  // namespace

size_t tiled_tiles_bytes(int64_t d, int64_t k) {
  const int TF = 2 * ks_of(d) * 32 + 32;
  return sizeof(float) * (size_t)(ceil_div<int64_t>(k, 32) * TF) + 1024;
}

namespace {

size_t fp32_lds_bytes(int64_t d, bool gather) {
  const int KS = ks_of(d);
  const int TF = 2 * KS * 32 + 32;
  int cpc = kChunkBytes / (4 * TF);
  if (cpc < 1) cpc = 1;
  const size_t buf = ((size_t)cpc * TF * 4 + 1023) / 1024 * 1024;
  return 2 * buf + (gather ? sizeof(int) * (kMaxSegments + 1) : 0);
}

// FP32 centroid tiles in the kernels' LDS layout into t
// (tiled_tiles_bytes(d, k) bytes: the tiles + a 1 KB read-ahead pad).
void prep_fp32_tiles_into(int64_t d, int64_t k, const float *C, const float *hn,
                          float *t, hipStream_t s) {
  const int KS = ks_of(d);
  const int TF = 2 * KS * 32 + 32;
  const int n_ct = (int)ceil_div<int64_t>(k, 32);
  const int64_t total = (int64_t)n_ct * TF;
  const dim3 grid((unsigned)ceil_div<int64_t>(total, 256));
  switch (KS) {
    case 16:
      hipLaunchKernelGGL(tiled_fp32_prep_krn<16>, grid, dim3(256), 0, s, (int)d,
                         (int)k, n_ct, C, hn, t);
      break;
    case 32:
      hipLaunchKernelGGL(tiled_fp32_prep_krn<32>, grid, dim3(256), 0, s, (int)d,
                         (int)k, n_ct, C, hn, t);
      break;
    default:
      hipLaunchKernelGGL(tiled_fp32_prep_krn<64>, grid, dim3(256), 0, s, (int)d,
                         (int)k, n_ct, C, hn, t);
  }
  KMD_LAUNCH_CHECK();
}

TempBuffer prep_fp32_tiles(int64_t d, int64_t k, const float *C, const float *hn,
                           hipStream_t s) {
  TempBuffer t(tiled_tiles_bytes(d, k), s);
  prep_fp32_tiles_into(d, k, C, hn, t.as<float>(), s);
  return t;
}

// T: BF16 terms (1, 3) or kF8 (the FP8 kernel)
template <typename I>
TiledMixedFn pick_mixed_tiled(int64_t d, int T) {
  const int D = bf16_d_of(d);
#define KMD_TM_CASE(DD)                                                     \
  if (D == DD)                                                              \
    return T == kF8 ? &tiled_fp8_krn<DD, I>                                 \
                    : (T == 3 ? &tiled_mixed_krn<DD, 3, I> : &tiled_mixed_krn<DD, 1, I>);
  KMD_TM_CASE(32)
  KMD_TM_CASE(64)
  KMD_TM_CASE(96)
  KMD_TM_CASE(128)
#undef KMD_TM_CASE
  return nullptr;
}

size_t mixed_tiled_lds_bytes(int64_t d, int T) {
  const int D = bf16_d_of(d);
  if (T == kF8)
    return 2 * (size_t)f8_cpc(D) * f8_slots(D) * 1024 + 16 + (size_t)f8_tml_bytes();
  return 2 * (size_t)mixed_cpc(D, T) * kk_of(D, T) * 64 * 16 + 16 +
         (size_t)mixed_tml_bytes(D, T);
}

// 32-sample blocks per wave of the BF16 (T = 1, 3) or FP8 (T = kF8) kernel
int mixed_s_of(int D, int T) { return T == kF8 ? f8_s() : bf16_s(D, T); }

// bound[0], bound[1] as tiled_bound_krn; the FP8 kernels' per-tile gathers
// bound[2] (eta / min s_t) and bound[5] (max s_t) start from 0
__global__ void __launch_bounds__(256)
    tiled_f8_bound_krn(int k, const float *__restrict__ hn, float *bound) {
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
    bound[2] = 0.f;
    bound[5] = 0.f;
  }
}

template <typename F>
int occupancy_grid(F fn, size_t lds) {
  static thread_local const void *attr_done[64] = {};
  const void *p = reinterpret_cast<const void *>(fn);
  bool seen = false;
  for (const void *f : attr_done) seen |= (f == p);
  if (!seen) {
    KMD_HIP_CHECK(hipFuncSetAttribute(p, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)kLdsPerCu));
    for (const void *&f : attr_done)
      if (f == nullptr) {
        f = p;
        break;
      }
  }
  int per_cu = 0;
  KMD_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, p, kTNT, lds));
  per_cu = apply_blocks_cap(per_cu);
  if (per_cu < 1) return 0;
  return device_cu_count() * per_cu;
}

inline int64_t update_blocks(int64_t n) {
  // chunks of >= 16K samples, at most 1024 blocks (each block claims its
  // share of every cluster with one global atomic)
  int64_t nb = ceil_div<int64_t>(n, 16384);
  if (nb > 1024) nb = 1024;
  return nb < 1 ? 1 : nb;
}

}  // namespace

bool tiled_supported(int64_t d, int64_t k) {
  // k: the update counts clusters in LDS
  return d >= 1 && d <= kMaxTiledD && k >= 1 && k <= 32768;
}

void launch_transpose(int64_t n, int64_t d, const float *X_t, float *X_r,
                      hipStream_t s) {
  if (n <= 0) return;
  const dim3 grid((unsigned)ceil_div<int64_t>(n, 32), (unsigned)ceil_div<int64_t>(d, 32));
  hipLaunchKernelGGL(transpose_krn, grid, dim3(256), 0, s, n, (int)d, X_t, X_r);
  KMD_LAUNCH_CHECK();
}

template <typename I>
void launch_tiled_assign(int64_t n, int64_t d, int64_t k, const float *X_t,
                         const float *C, const float *hn, I *labels,
                         hipStream_t s, const int *skip, void *tiles_scratch) {
  if (n <= 0) return;
  require(tiled_supported(d, k), "tiled assignment: unsupported shape");
  TiledFn fn = pick_fp32<I, false>(d);
  const size_t lds = fp32_lds_bytes(d, false);
  const int grid_max = occupancy_grid(fn, lds);
  if (grid_max < 1) throw RuntimeError("tiled assignment: no launch plan");
  const int S = ks_of(d) <= 32 ? 2 : 1;
  const int64_t need = ceil_div<int64_t>(ceil_div<int64_t>(n, 32 * S), kTW);
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(grid_max, need));
  TempBuffer own(tiles_scratch ? 0 : tiled_tiles_bytes(d, k), s);
  float *tiles = tiles_scratch ? static_cast<float *>(tiles_scratch) : own.as<float>();
  prep_fp32_tiles_into(d, k, C, hn, tiles, s);
  TiledArgs a{};
  a.n = n;
  a.d = (int)d;
  a.k = (int)k;
  a.n_ct = (int)ceil_div<int64_t>(k, 32);
  a.X = X_t;
  a.tiles = tiles;
  a.labels = labels;
  a.skip = skip;
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(kTNT), lds, s, a);
  KMD_LAUNCH_CHECK();
}

namespace {
template <typename I>
void launch_rescore_full(int64_t n, int64_t d, int64_t k, const float *X_r,
                         const float *tiles, I *labels, const int *queue,
                         const int *queue_counts, int n_segments, int64_t seg_cap,
                         int *n_rescored, const int *only_if, hipStream_t s) {
  require(n_segments >= 1 && n_segments <= kMaxSegments,
          "tiled re-score: segment count out of range");
  TiledFn fn = pick_fp32<I, true>(d);
  const size_t lds = fp32_lds_bytes(d, true);
  const int grid = occupancy_grid(fn, lds);
  if (grid < 1) throw RuntimeError("tiled re-score: no launch plan");
  TiledArgs a{};
  a.n = n;
  a.d = (int)d;
  a.k = (int)k;
  a.n_ct = (int)ceil_div<int64_t>(k, 32);
  a.X = X_r;
  a.tiles = tiles;
  a.labels = labels;
  a.queue = queue;
  a.queue_counts = queue_counts;
  a.n_segments = n_segments;
  a.seg_cap = seg_cap;
  a.n_rescored = n_rescored;
  a.only_if = only_if;
  hipLaunchKernelGGL(fn, dim3(grid), dim3(kTNT), lds, s, a);
  KMD_LAUNCH_CHECK();
}
}  // namespace

template <typename I>
void launch_tiled_rescore(int64_t n, int64_t d, int64_t k, const float *X_r,
                          const float *C, const float *hn, I *labels,
                          const int *queue, const int *queue_counts,
                          int n_segments, int64_t seg_cap, int *n_rescored,
                          hipStream_t s) {
  TempBuffer tiles = prep_fp32_tiles(d, k, C, hn, s);
  launch_rescore_full<I>(n, d, k, X_r, tiles.as<const float>(), labels, queue,
                         queue_counts, n_segments, seg_cap, n_rescored, nullptr, s);
}

bool tiled_cand_supported(int64_t k) { return ceil_div<int64_t>(k, 32) <= 32; }

size_t tiled_cand_scratch_bytes(int64_t n, int64_t k) {
  const int64_t n_ct = ceil_div<int64_t>(k, 32);
  return sizeof(int) * (size_t)(4 * (n_ct + 1) + 4) +
         sizeof(int) * (size_t)(8 * (n > 0 ? n : 1)) +
         sizeof(unsigned long long) * (size_t)(n > 0 ? n : 1) + 512;
}

template <typename I>
void launch_tiled_cand_rescore(int64_t n, int64_t d, int64_t k, const float *X_r,
                               const float *C, const float *hn, I *labels,
                               const int *queue, const int *queue_counts,
                               const unsigned *qmask, int n_segments,
                               int64_t seg_cap, void *scratch, int *n_rescored,
                               hipStream_t s) {
  require(tiled_cand_supported(k), "candidate re-score: more than 32 tiles");
  const int n_ct = (int)ceil_div<int64_t>(k, 32);
  TempBuffer tiles = prep_fp32_tiles(d, k, C, hn, s);
  CandArgs a{};
  a.n_segments = n_segments;
  a.n_ct = n_ct;
  a.d = (int)d;
  a.seg_cap = seg_cap;
  a.queue = queue;
  a.queue_counts = queue_counts;
  a.qmask = qmask;
  int *ip = reinterpret_cast<int *>(scratch);
  a.tile_cnt = ip;
  a.tile_cursor = ip + (n_ct + 1);
  a.tile_start = ip + 2 * (n_ct + 1);
  a.tile_witem = ip + 3 * (n_ct + 1);
  a.overflow = ip + 4 * (n_ct + 1);
  a.pairs = ip + 4 * (n_ct + 1) + 4;
  a.pair_cap = 8 * (n > 0 ? n : 1);
  a.keys = reinterpret_cast<unsigned long long *>(
      (reinterpret_cast<uintptr_t>(a.pairs + a.pair_cap) + 255) & ~uintptr_t(255));
  a.Xr = X_r;
  a.tiles = tiles.as<const float>();
  a.labels = labels;
  a.n_rescored = n_rescored;
  hipLaunchKernelGGL(fill_zero_i32_krn, dim3(1), dim3(256), 0, s, 2 * (n_ct + 1),
                     a.tile_cnt);
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(cand_count_krn, dim3(n_segments), dim3(256), 0, s, a);
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(cand_scan_krn, dim3(1), dim3(64), 0, s, a);
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(cand_scatter_krn, dim3(n_segments), dim3(256), 0, s, a);
  KMD_LAUNCH_CHECK();
  const int grid = device_cu_count() * 2;
  switch (ks_of(d)) {
    case 16:
      hipLaunchKernelGGL(cand_rescore_krn<16>, dim3(grid), dim3(kTNT), 0, s, a);
      break;
    case 32:
      hipLaunchKernelGGL(cand_rescore_krn<32>, dim3(grid), dim3(kTNT), 0, s, a);
      break;
    default:
      hipLaunchKernelGGL(cand_rescore_krn<64>, dim3(grid), dim3(kTNT), 0, s, a);
  }
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL((cand_finalize_krn<I>), dim3(n_segments), dim3(256), 0, s, a);
  KMD_LAUNCH_CHECK();
  // the pairs did not fit: the full gather re-score of every queued sample
  launch_rescore_full<I>(n, d, k, X_r, tiles.as<const float>(), labels, queue,
                         queue_counts, n_segments, seg_cap, n_rescored, a.overflow, s);
}

template <typename I>
int tiled_mixed_blocks(int64_t n, int64_t d, int64_t k, int T) {
  if (!tiled_supported(d, k) || n <= 0 || n >= (int64_t(1) << 31)) return 0;
  if (mixed_idx_bits(k) > 12) return 0;  // packed index: <= 4096 centroids
  TiledMixedFn fn = pick_mixed_tiled<I>(d, T);
  if (fn == nullptr) return 0;
  const int grid = occupancy_grid(fn, mixed_tiled_lds_bytes(d, T));
  if (grid < 1) return 0;
  const int S = mixed_s_of(bf16_d_of(d), T);
  const int64_t need = ceil_div<int64_t>(ceil_div<int64_t>(n, 32 * S), kTW);
  int64_t blocks = std::min<int64_t>(grid, need);
  if (blocks > kMaxSegments) blocks = kMaxSegments;
  return (int)std::max<int64_t>(1, blocks);
}

int64_t tiled_mixed_seg_cap(int64_t n, int64_t d, int T, int blocks) {
  const int S = mixed_s_of(bf16_d_of(d), T);
  const int64_t n_super = ceil_div<int64_t>(n > 0 ? n : 1, 32 * S);
  const int64_t rounds = ceil_div<int64_t>(n_super, (int64_t)blocks * kTW);
  return rounds * kTW * 32 * S;
}

size_t tiled_xb_bytes(int64_t n, int64_t d) {
  return sizeof(unsigned short) * (size_t)(n > 0 ? n : 1) * bf16_d_of(d) + 256;
}

void launch_tiled_xb_prep(int64_t n, int64_t d, const float *X_t, void *xb,
                          float *xnb, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(tiled_xb_prep_krn, dim3((unsigned)ceil_div<int64_t>(n, 256)),
                     dim3(256), 0, s, n, (int)d, bf16_d_of(d), X_t,
                     reinterpret_cast<bf16x8 *>(xb), xnb);
  KMD_LAUNCH_CHECK();
}

size_t tiled_aop_bytes(int64_t d, int64_t k, int T) {
  // + 1 KB: the chunk copies read whole 1 KB pieces
  if (T == kF8)
    return (size_t)ceil_div<int64_t>(k, 32) * f8_slots(bf16_d_of(d)) * 1024 + 1024;
  return (size_t)ceil_div<int64_t>(k, 32) * kk_of(bf16_d_of(d), T) * 64 * 16 + 1024;
}

size_t tiled_bound_floats() { return 8; }

void launch_tiled_x8_prep(int64_t n, int64_t d, const float *X_t, void *x8,
                          float *xnb, float *bound, hipStream_t s) {
  if (n <= 0) return;
  TempBuffer mb(sizeof(unsigned), s);
  KMD_HIP_CHECK(hipMemsetAsync(mb.as<unsigned>(), 0, sizeof(unsigned), s));
  const int64_t count = n * d;
  const int grid = (int)std::max<int64_t>(
      1, std::min<int64_t>(ceil_div<int64_t>(count, 256 * 16), 4 * device_cu_count()));
  hipLaunchKernelGGL(tiled_x8_max_krn, dim3(grid), dim3(256), 0, s, count, X_t,
                     mb.as<unsigned>());
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(tiled_x8_prep_krn, dim3((unsigned)ceil_div<int64_t>(n, 256)),
                     dim3(256), 0, s, n, (int)d, bf16_d_of(d), X_t,
                     mb.as<const unsigned>(), reinterpret_cast<unsigned char *>(x8), xnb,
                     bound);
  KMD_LAUNCH_CHECK();
}

void launch_f8_roundtrip(int64_t n, const float *in, float *out, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(f8_roundtrip_krn, dim3((unsigned)ceil_div<int64_t>(n, 256)),
                     dim3(256), 0, s, n, in, out);
  KMD_LAUNCH_CHECK();
}

void launch_tiled_mixed_prep(int64_t d, int64_t k, int T, const float *C,
                             const float *hn, void *aop, float *bound,
                             hipStream_t s) {
  const int D = bf16_d_of(d);
  const int n_ct = (int)ceil_div<int64_t>(k, 32);
  if (T == kF8) {
    hipLaunchKernelGGL(tiled_f8_bound_krn, dim3(1), dim3(256), 0, s, (int)k, hn, bound);
    KMD_LAUNCH_CHECK();
    hipLaunchKernelGGL(tiled_fp8_prep_krn, dim3((unsigned)n_ct), dim3(256), 0, s,
                       (int)d, (int)k, D, C, hn, reinterpret_cast<unsigned char *>(aop),
                       bound);
    KMD_LAUNCH_CHECK();
    return;
  }
  const int KK = kk_of(D, T);
  const int64_t total = (int64_t)n_ct * KK * 64;
  hipLaunchKernelGGL(tiled_prep_krn, dim3((unsigned)ceil_div<int64_t>(total, 256)),
                     dim3(256), 0, s, (int)d, (int)k, D, T, KK, n_ct, C, hn,
                     reinterpret_cast<bf16x8 *>(aop));
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(tiled_bound_krn, dim3(1), dim3(256), 0, s, (int)k, hn, bound);
  KMD_LAUNCH_CHECK();
}

template <typename I>
void launch_tiled_mixed_assign(int64_t n, int64_t d, int64_t k, int T,
                               int blocks, const float *X_t, const void *aop,
                               const float *bound, I *labels, int *queue,
                               int *queue_counts, unsigned *qmask, hipStream_t s,
                               const int *skip, const void *xb, const float *xnb) {
  TiledMixedFn fn = pick_mixed_tiled<I>(d, T);
  if (fn == nullptr || blocks < 1)
    throw RuntimeError("tiled BF16 assignment: unsupported shape");
  TiledMixedArgs a{};
  a.n = n;
  a.d = (int)d;
  a.k = (int)k;
  a.n_ct = (int)ceil_div<int64_t>(k, 32);
  a.idx_bits = mixed_idx_bits(k);
  a.fp32_len = 2 * ks_of(d);
  a.X = X_t;
  a.aop = reinterpret_cast<const bf16x8 *>(aop);
  a.bound = bound;
  a.labels = labels;
  a.queue = queue;
  a.queue_counts = queue_counts;
  a.qmask = tiled_cand_supported(k) ? qmask : nullptr;
  a.seg_cap = tiled_mixed_seg_cap(n, d, T, blocks);
  a.pinf = std::numeric_limits<float>::infinity();
  a.skip = skip;
  if ((T == 1 || T == kF8) && (xb == nullptr || xnb == nullptr))
    throw RuntimeError("tiled BF16 / FP8 assignment: the sample copy is required");
  a.xb = (T == 1 || T == kF8) ? reinterpret_cast<const bf16x8 *>(xb) : nullptr;
  a.xnb = (T == 1 || T == kF8) ? xnb : nullptr;
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(kTNT), mixed_tiled_lds_bytes(d, T), s,
                     a);
  KMD_LAUNCH_CHECK();
}

size_t tiled_update_scratch_bytes(int64_t n, int64_t k) {
  return sizeof(int) * (size_t)(4 * (k + 1) + (n > 0 ? n : 1)) + 256;
}

template <typename I>
void launch_tiled_update(int64_t n, int64_t d, int64_t k, const float *X_r,
                         const float *w, const I *labels, const float *scale,
                         void *scratch, void *sums, hipStream_t s,
                         const int *skip) {
  const int64_t nb = update_blocks(n);
  const int64_t chunk = ceil_div<int64_t>(n > 0 ? n : 1, nb);
  int *totals = reinterpret_cast<int *>(scratch);  // [k + 1]
  int *cursor = totals + (k + 1);                  // [k + 1]
  int *start = cursor + (k + 1);                   // [k + 1]
  int *items = start + (k + 1);                    // [k + 1]
  int *perm = items + (k + 1);                     // [n]
  u64 *S = reinterpret_cast<u64 *>(sums);
  const int64_t m = (d + 1) * k;
  hipLaunchKernelGGL(fill_u64_krn,
                     dim3((unsigned)std::min<int64_t>(1024, ceil_div<int64_t>(m, 256))),
                     dim3(256), 0, s, m, S);
  KMD_LAUNCH_CHECK();
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_i32_krn, dim3(8), dim3(256), 0, s, 2 * (k + 1), totals);
  KMD_LAUNCH_CHECK();
  const size_t hist_lds = sizeof(int) * (size_t)k;
  require(hist_lds <= kLdsPerCu, "tiled update: too many clusters for LDS counts");
  hipLaunchKernelGGL((label_hist_krn<I>), dim3((unsigned)nb), dim3(256), hist_lds, s,
                     n, (int)k, chunk, labels, totals, skip);
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(cluster_scan_krn, dim3(1), dim3(1024), 0, s, (int)k, kMembers,
                     totals, start, items, skip);
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL((scatter_krn<I>), dim3((unsigned)nb), dim3(256), hist_lds, s, n,
                     (int)k, chunk, labels, start, cursor, perm, skip);
  KMD_LAUNCH_CHECK();
  // at most n / kMembers + k work items; blocks beyond items[k] return
  const int64_t max_items = n / kMembers + k;
  const int threads =
      (int)std::min<int64_t>(320, ceil_div<int64_t>(d + 1, kWave) * kWave);
  hipLaunchKernelGGL(segment_sum_krn, dim3((unsigned)max_items), dim3(threads), 0, s,
                     (int)d, (int)k, X_r, w, scale, perm, start, items, S, skip);
  KMD_LAUNCH_CHECK();
}

// ---- the reference fused step's update on the tiled shapes -----------------

// copies[b % n_copies][f][l_s] += w_s x_{f,s}, sizes_copies[b % n_copies][l_s]
// += w_s (src/lloyd_single_step.hpp:170-204, the accumulate-into contract).
// Block (b, g) takes samples [b chunk, (b + 1) chunk) and feature group g of
// F features (g == n_groups - 1: the counts); the group's totals live in LDS
// as FP32 (ds_add_f32, order-dependent like the reference's float atomics)
// and go to the copy with one global atomic per non-zero cell.  X rows are
// read coalesced; labels and weights once per group.
template <typename I>
__global__ void __launch_bounds__(512)
    label_update_krn(int64_t n, int d, int k, int F, int64_t chunk,
                     const float *__restrict__ X, const float *__restrict__ w,
                     const I *__restrict__ labels, float *copies,
                     float *sizes_copies, int n_copies) {
  extern __shared__ float acc[];
  const int g = blockIdx.y;
  const bool counts = g == (int)gridDim.y - 1;
  const int f0 = g * F;
  const int nf = counts ? 1 : min(F, d - f0);
  const int cells = nf * k;
  for (int i = threadIdx.x; i < cells; i += blockDim.x) acc[i] = 0.f;
  __syncthreads();
  const int64_t s0 = (int64_t)blockIdx.x * chunk;
  const int64_t s1 = min(n, s0 + chunk);
  constexpr int kF = 8;   // feature loads in flight per thread
  for (int64_t s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
    const int l = (int)labels[s];
    const float ws = w[s];
    if (counts) {
      atomicAdd(&acc[l], ws);
      continue;
    }
    const float *col = X + (int64_t)f0 * n + s;
    for (int f = 0; f < nf; f += kF) {
      float x[kF];
#pragma unroll
      for (int j = 0; j < kF; ++j) x[j] = f + j < nf ? col[(int64_t)(f + j) * n] : 0.f;
#pragma unroll
      for (int j = 0; j < kF; ++j)
        if (f + j < nf) atomicAdd(&acc[(f + j) * k + l], ws * x[j]);
    }
  }
  __syncthreads();
  const int copy = (int)(blockIdx.x % (unsigned)n_copies);
  float *dst = counts ? sizes_copies + (int64_t)copy * k
                      : copies + ((int64_t)copy * d + f0) * k;
  for (int i = threadIdx.x; i < cells; i += blockDim.x) {
    const float v = acc[i];
    if (v != 0.f) atomicAdd(&dst[i], v);
  }
}

template <typename I>
void launch_label_update(int64_t n, int64_t d, int64_t k, int64_t n_copies,
                         const float *X_t, const float *w, const I *labels,
                         float *copies, float *sizes_copies, hipStream_t s) {
  if (n <= 0) return;
  require(k <= 40960, "label update: too many clusters for one LDS row");
  // F features per group: F k floats <= 64 KB (two blocks per CU)
  const int F = (int)std::max<int64_t>(1, std::min<int64_t>(d, 16384 / k));
  const int groups = (int)ceil_div<int64_t>(d, F) + 1;
  const size_t lds = sizeof(float) * (size_t)F * k;
  const int per_cu = lds <= 80 * 1024 ? 2 : 1;
  const int64_t target = (int64_t)device_cu_count() * per_cu * 2;
  const int64_t b = std::max<int64_t>(
      1, std::min<int64_t>(ceil_div<int64_t>(target, groups), ceil_div<int64_t>(n, 512)));
  const int64_t chunk = ceil_div<int64_t>(n, b);
  if (lds > 64 * 1024) allow_full_lds(reinterpret_cast<const void *>(&label_update_krn<I>));
  hipLaunchKernelGGL(label_update_krn<I>, dim3((unsigned)ceil_div<int64_t>(n, chunk),
                                              (unsigned)groups),
                     dim3(512), lds, s, n, (int)d, (int)k, F, chunk, X_t, w, labels,
                     copies, sizes_copies, (int)n_copies);
  KMD_LAUNCH_CHECK();
}

#define KMD_INST_TILED(I)                                                      \
  template void launch_tiled_assign<I>(int64_t, int64_t, int64_t,              \
                                       const float *, const float *,           \
                                       const float *, I *, hipStream_t,        \
                                       const int *, void *);                   \
  template void launch_label_update<I>(int64_t, int64_t, int64_t, int64_t,     \
                                       const float *, const float *,           \
                                       const I *, float *, float *,            \
                                       hipStream_t);                           \
  template void launch_tiled_rescore<I>(int64_t, int64_t, int64_t,             \
                                        const float *, const float *,          \
                                        const float *, I *, const int *,       \
                                        const int *, int, int64_t, int *,      \
                                        hipStream_t);                          \
  template int tiled_mixed_blocks<I>(int64_t, int64_t, int64_t, int);          \
  template void launch_tiled_mixed_assign<I>(int64_t, int64_t, int64_t, int,   \
                                             int, const float *, const void *, \
                                             const float *, I *, int *, int *, \
                                             unsigned *, hipStream_t,          \
                                             const int *, const void *,        \
                                             const float *);                   \
  template void launch_tiled_cand_rescore<I>(                                  \
      int64_t, int64_t, int64_t, const float *, const float *, const float *,  \
      I *, const int *, const int *, const unsigned *, int, int64_t, void *,   \
      int *, hipStream_t);                                                     \
  template void launch_tiled_update<I>(int64_t, int64_t, int64_t,              \
                                       const float *, const float *,           \
                                       const I *, const float *, void *,       \
                                       void *, hipStream_t, const int *);
KMD_INST_TILED(int32_t)
KMD_INST_TILED(int64_t)

}  // namespace kmd
