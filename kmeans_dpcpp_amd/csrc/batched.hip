// Many independent Lloyd fits, one launch per step of an iteration for all
// of them, the iteration captured once in a hipGraph and replayed.
//
// Work lists.  The host flattens the fits into (fit, run of samples) items
// for the assignment (one list per feature bucket d <= 16 / 32 / 64, one
// persistent launch each; heaviest items first) and (fit, 4096-sample chunk)
// items for the update.
// A block takes items grid-stride, loads the item's fit's centroid tiles and
// half norms into LDS and labels the tile's samples with the FP32 MFMA
// arithmetic of the single-fit kernels (v_mfma_f32_32x32x2f32 over ascending
// k-steps, argmin of hn - dot, strict <, lowest index on ties), one 32-sample
// column block per wave.  Items of a fit whose mask is off are skipped, so
// converged fits cost nothing.
//
// Update.  Per fit, exactly as the tiled single-fit path: labels counted,
// cluster totals scanned, samples scattered into cluster order and summed
// as int64 fixed point (the fit's own scales) from a sample-major copy of
// its X, one block per segment item (up to 256 members of one cluster; the
// items of all fits are flattened by a map kernel).  Integer sums do not
// depend on the order, so they equal the single-fit driver's bit for bit.
//
// Finalize (one block per fit): sums -> centroids and sizes, ascending
// empty-cluster list, relocation of empty clusters to the farthest samples
// (sq-dist desc, index asc), division, centroid shifts, the shift sum in the
// same FP64 order as the single-fit reduction, then it++ and the fit's mask
// (it < max_iter && shift > tol).  With the same labels, sums and reduction
// orders, every fit ends with the single-fit driver's labels, centroids,
// iteration count and inertia.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "common.hpp"
#include "fixed_point.hpp"
#include "lloyd_fast.hpp"

namespace kmd {

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned long long u64;

// Assignment items are sized by their MFMA work: about kItemMacs
// multiply-adds (samples x centroids padded to 32-tiles x features padded to
// the kernel's feature bucket), at least kItemMin and at most kItemMax
// samples, in multiples of kItemMin (one sub-tile of either block shape).
constexpr int64_t kItemMacs = int64_t(2048) * 256 * 64;
constexpr int kItemMin = 512, kItemMax = 2048;
// assignment block: one 32-sample column block per wave; 16 waves, or 8 for
// d > 32, whose 32 x-registers per lane do not fit the 128 VGPRs of 16 waves
constexpr int assign_threads(int KS) { return KS <= 16 ? 1024 : 512; }
                                          // (the fit's centroids are loaded into
                                          // LDS once per item)
constexpr int kChunk = 4096;              // samples per update item
constexpr int kPerThread = kChunk / 256;  // of a 256-thread update block
static_assert(kChunk % 256 == 0, "update items split evenly over 256 threads");
constexpr int kMembers = 256;             // members per segment-sum item
constexpr int kMaxK = 256;
constexpr int kMaxD = 64;
constexpr int kFB = 256;                  // finalize / per-fit block

struct BState {
  int it, active, cur, n_empty;
  int seg_items, seg_base;
  double shift;
  double inertia;
};

// One fit: the caller's arrays and the fit's own device workspace (so fits
// can join and leave a running set without moving anyone's state).
struct BFit {
  int64_t n;
  int d, k;
  const float *X, *w;
  float *C[2];      // C[0] = init (overwritten), C[1] = result
  int *labels;
  int64_t max_iter;
  float tol;
  BState *st;
  float *xr;        // (n, d) sample-major copy of X
  int *perm;        // (n) samples in cluster order
  float *dist;      // (n) relocation distances / per-sample inertia
  u64 *sums;        // ((d + 1) k) int64 fixed-point sums
  float *scale;     // (d + 1) fixed-point scales
  float *hn, *sizes;                        // (k + 1)
  int *list, *totals, *cursor, *start, *citems;  // (k + 1)
  int *nonfinite;   // (1) 0: fixed-point sums; 1: inf/NaN; 2: range too wide
  unsigned *maxbits;  // (d + 2) range maxima (fixed_point.hpp)
  u64 *estat;         // (2 (d + 1)) exponent sums and counts (fixed_point.hpp)
};

// A pointer read from a BFit, marked as global memory: loads through it are
// global_load (counted by vmcnt only), not flat loads, whose lgkmcnt share
// makes every LDS wait also wait for them.
template <typename T>
__device__ inline const __attribute__((address_space(1))) T *gptr(const T *p) {
  return (const __attribute__((address_space(1))) T *)p;
}

// The current (c = 0) or other centroid buffer.  A select, not F.C[c]: a
// dynamic index into a struct held in registers puts the struct in scratch.
__device__ inline float *cbuf(const BFit &F, int c) { return c ? F.C[1] : F.C[0]; }

struct AItem {  // assignment tile: samples first .. first + count - 1 of a fit
  int fit;
  int first, count;
};
struct CItem {  // update chunk
  int fit;
  int first, count;
};
struct TItem {  // transpose tile: 32 samples x 32 features of one fit
  int fit;
  int s0, f0;
};

__device__ inline double block_sum256(double v, double *lds) {
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  const int w = threadIdx.x / kWave;
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) lds[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < (int)(blockDim.x / kWave); ++i) t += lds[i];
  return t;
}

// ---- half norms -------------------------------------------------------------
__global__ void __launch_bounds__(kFB)
    bt_hn_krn(const BFit *__restrict__ fits, int all) {
  const int f = blockIdx.x;
  if (!all && !fits[f].st[0].active) return;
  const BFit F = fits[f];
  const float *C = cbuf(F, fits[f].st[0].cur);
  constexpr int kU = 8;  // loads in flight per thread; the fmaf order is j = 0, 1, ...
  for (int c = threadIdx.x; c < F.k; c += kFB) {
    float acc = 0.f;
    for (int j0 = 0; j0 < F.d; j0 += kU) {
      float v[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u)
        v[u] = (j0 + u < F.d) ? C[(int64_t)(j0 + u) * F.k + c] : 0.f;
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (j0 + u < F.d) acc = fmaf(v[u], v[u], acc);
    }
    F.hn[c] = acc / 2.f;
  }
}

// ---- assignment -----------------------------------------------------------------
template <int KS>
constexpr int tile_floats() {
  return 2 * KS * 32 + 32;
}

template <int KS, int NT = assign_threads(KS)>
__global__ void __launch_bounds__(NT)
    bt_assign_krn(const BFit *__restrict__ fits, const AItem *__restrict__ items, int n_items,
                  int all) {
  constexpr int TF = tile_floats<KS>();
  constexpr int kNT = NT, kW = kNT / kWave;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  float *Cl = reinterpret_cast<float *>(lds_raw);  // [kMaxK / 32][TF]
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1), wave = tid / kWave;
  const int col = lane & 31, half = lane >> 5;
  // The block's items are blockIdx.x, + gridDim.x, ... (grid-stride); the
  // mask flags of the next 64 of them are read at once by wave 0 and the
  // active ones listed in LDS, so converged fits cost one flag read per item
  // in parallel instead of a round trip each.
  __shared__ int act[kWave];
  __shared__ int n_act;
  for (int base = blockIdx.x; base < n_items; base += kWave * gridDim.x) {
    __syncthreads();  // the previous batch's list is no longer read
    if (wave == 0) {
      const int it = base + lane * gridDim.x;
      const bool a = it < n_items && (all || fits[items[it].fit].st[0].active);
      const unsigned long long m = __ballot(a);
      if (a) act[__popcll(m & ((1ull << lane) - 1))] = it;
      if (lane == 0) n_act = __popcll(m);
    }
    __syncthreads();
    const int na = n_act;
    for (int j = 0; j < na; ++j) {
      const AItem item = items[act[j]];
      const BFit F = fits[item.fit];
      const auto *C = gptr(cbuf(F, fits[item.fit].st[0].cur));
      const auto *hn = gptr(F.hn);
      const int d = F.d, k = F.k, n_ct = (k + 31) / 32;
      __syncthreads();
      // stage the centroid tiles: four loads in flight per thread, then the stores
      constexpr int kSU = 4;
      for (int i0 = tid; i0 < n_ct * TF; i0 += kSU * kNT) {
        float v[kSU];
        int at[kSU];
#pragma unroll
        for (int u = 0; u < kSU; ++u) {
          const int i = i0 + u * kNT;
          const int t = i / TF, e = i - t * TF;
          at[u] = -1;
          v[u] = 0.f;
          if (i < n_ct * TF) {
            if (e < 2 * KS * 32) {
              const int f = e >> 5, cc = e & 31;
              const int c = t * 32 + cc;
              const int ks = f >> 1, hf = f & 1;
              at[u] = t * TF + ((ks >> 2) * 2 + hf) * 128 + cc * 4 + (ks & 3);
              v[u] = (f < d && c < k) ? C[(int64_t)f * k + c] : 0.f;
            } else {
              const int c = t * 32 + (e - 2 * KS * 32);
              at[u] = t * TF + e;
              v[u] = (c < k) ? -hn[c] : -__builtin_huge_valf();  // -hn
            }
          }
        }
#pragma unroll
        for (int u = 0; u < kSU; ++u)
          if (at[u] >= 0) Cl[at[u]] = v[u];
      }
      __syncthreads();
      const int64_t n = F.n;
      const int64_t end = min((int64_t)item.first + item.count, n);
      for (int64_t s0 = item.first; s0 < end; s0 += kW * 32) {  // block-uniform
        const int64_t p = s0 + wave * 32 + col;
        float xr[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const int f = 2 * ks + half;
          xr[ks] = (f < d && p < end) ? gptr(F.X)[(int64_t)f * n + p] : 0.f;
        }
        float bestv = __builtin_huge_valf();
        int besti = 0;
        for (int t = 0; t < n_ct; ++t) {
          const float *At = Cl + t * TF;
          // the chain starts from the tile's -hn, so -acc = hn - c.x
          const float *Ht = At + 2 * KS * 32;
          f32x16 acc;
#pragma unroll
          for (int gg = 0; gg < 4; ++gg) {
            const f32x4 h4 = *reinterpret_cast<const f32x4 *>(Ht + 8 * gg + 4 * half);
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[4 * gg + e] = h4[e];
          }
#pragma unroll
          for (int k4 = 0; k4 < KS / 4; ++k4) {
            const f32x4 av =
                *reinterpret_cast<const f32x4 *>(At + (k4 * 2 + half) * 128 + col * 4);
#pragma unroll
            for (int j = 0; j < 4; ++j)
              acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[j], xr[4 * k4 + j], acc, 0, 0, 0);
          }
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
        }
        const float ov = __shfl_xor(bestv, 32, kWave);
        const int oi = __shfl_xor(besti, 32, kWave);
        const bool take = (ov < bestv) || (ov == bestv && oi < besti);
        besti = take ? oi : besti;
        if (half == 0 && p < end) F.labels[p] = besti;
      }
    }
  }
}


// ---- assignment with the BF16 cross term and a certified FP32 re-score ----------
//
// The same items, staging and labels as bt_assign_krn, with the cross term of
// each (sample, centroid) pair first taken on the matrix core in BF16:
//   D = sum_f bf16(x_f) bf16(c_f) - (h1 + h2 + h3)   (v_mfma_f32_32x32x16_bf16,
// FP32 accumulation; h1 + h2 + h3 = the FP32 half norm split exactly into
// three BF16 terms, the samples' side of those K columns is -1), so that the
// argmin of hn - c.x is the argmax of D.  Each lane keeps the best and
// second-best D of its sample with the centroid index packed into the low
// mantissa bits.  With E the bound on the error of D against exact
// arithmetic plus the bound on the FP32 path's error (from ||x||,
// max ||c||, max hn, the index packing), a sample whose gap best - second
// exceeds 2 E has the FP32 path's label, unique, and is labelled here.  The
// others are listed in LDS and re-scored after the item's samples with the
// FP32 path's instruction sequence (bt_assign_krn's chain), so every label is
// the FP32 path's label bit for bit and the exact update that follows gives
// bitwise the FP32 fit.
template <int KS>
constexpr int mixed_kk() {
  return (2 * KS + 3 + 15) / 16;  // k-steps of 16: D = 2 KS features + 3 hn terms
}
template <int KS>
constexpr size_t mixed_aop_offset() {
  return (size_t)(kMaxK / 32) * (2 * KS * 32 + 32) * 4;   // after the FP32 tiles
}
template <int KS>
constexpr size_t mixed_list_offset() {
  return mixed_aop_offset<KS>() + (size_t)(kMaxK / 32) * mixed_kk<KS>() * 64 * 16;
}
template <int KS>
constexpr size_t mixed_assign_lds() {
  return mixed_list_offset<KS>() + sizeof(int) * kItemMax;
}

// v_max3_f32 on values known to be finite (no NaN canonicalisation in front)
__device__ inline float max3f(float a, float b, float c) {
  float r;
  asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

template <int KS, int NT = assign_threads(KS)>
__global__ void __launch_bounds__(NT)
    bt_assign_mixed_krn(const BFit *__restrict__ fits, const AItem *__restrict__ items,
                        int n_items, int all, float pinf) {
  constexpr int TF = tile_floats<KS>();
  constexpr int D = 2 * KS, KD = D / 16, KK = mixed_kk<KS>(), XH = D / 2;
  constexpr int kNT = NT, kW = kNT / kWave;
  // a lane's row r (0..15) of a tile is packed into the 4 low mantissa bits
  constexpr int kRowBits = 4;
  constexpr unsigned kRowMask = (1u << kRowBits) - 1u;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  float *Cl = reinterpret_cast<float *>(lds_raw);  // FP32 tiles [kMaxK / 32][TF]
  bf16x8 *Al = reinterpret_cast<bf16x8 *>(lds_raw + mixed_aop_offset<KS>());
  int *list = reinterpret_cast<int *>(lds_raw + mixed_list_offset<KS>());
  __shared__ int act[kWave];
  __shared__ int n_act, n_list;
  __shared__ float red[kW];
  __shared__ int bad_hn[kW];
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1), wave = tid / kWave;
  const int col = lane & 31, half = lane >> 5;
  constexpr float u = 1.0f / 256.0f;   // bf16 unit roundoff
  constexpr float acc_err = (float)(D + 3 + D + 8) * 5.9604645e-8f;
  constexpr float alpha = 2.f * u + u * u + acc_err;
  constexpr float beta = acc_err;
  const float pack_err = ldexpf(1.0f, kRowBits - 23) * 1.01f;
  for (int base = blockIdx.x; base < n_items; base += kWave * gridDim.x) {
    __syncthreads();
    if (wave == 0) {
      const int it = base + lane * gridDim.x;
      const bool a = it < n_items && (all || fits[items[it].fit].st[0].active);
      const unsigned long long m = __ballot(a);
      if (a) act[__popcll(m & ((1ull << lane) - 1))] = it;
      if (lane == 0) n_act = __popcll(m);
    }
    __syncthreads();
    const int na = n_act;
    for (int j = 0; j < na; ++j) {
      const AItem item = items[act[j]];
      const BFit F = fits[item.fit];
      const auto *C = gptr(cbuf(F, fits[item.fit].st[0].cur));
      const auto *hn = gptr(F.hn);
      const int d = F.d, k = F.k, n_ct = (k + 31) / 32;
      __syncthreads();
      if (tid == 0) n_list = 0;
      // FP32 tiles (the re-score's operands), as bt_assign_krn stages them,
      // and max hn over the fit's centroids
      float hmax = 0.f;
      int hbad = 0;
      constexpr int kSU = 4;
      for (int i0 = tid; i0 < n_ct * TF; i0 += kSU * kNT) {
        float v[kSU];
        int at[kSU];
#pragma unroll
        for (int uu = 0; uu < kSU; ++uu) {
          const int i = i0 + uu * kNT;
          const int t = i / TF, e = i - t * TF;
          at[uu] = -1;
          v[uu] = 0.f;
          if (i < n_ct * TF) {
            if (e < 2 * KS * 32) {
              const int f = e >> 5, cc = e & 31;
              const int c = t * 32 + cc;
              const int ks = f >> 1, hf = f & 1;
              at[uu] = t * TF + ((ks >> 2) * 2 + hf) * 128 + cc * 4 + (ks & 3);
              v[uu] = (f < d && c < k) ? C[(int64_t)f * k + c] : 0.f;
            } else {
              const int c = t * 32 + (e - 2 * KS * 32);
              at[uu] = t * TF + e;
              if (c < k) {
                const float h = hn[c];
                v[uu] = -h;
                hbad |= !(h < 1e36f);         // also true for NaN
                hmax = fmaxf(hmax, h);
              } else {
                v[uu] = -__builtin_huge_valf();
              }
            }
          }
        }
#pragma unroll
        for (int uu = 0; uu < kSU; ++uu)
          if (at[uu] >= 0) Cl[at[uu]] = v[uu];
      }
      // BF16 A operands: element jj of lane l for (tile ct, k-step kk) is
      // centroid ct 32 + l % 32 at K position 16 kk + 8 (l / 32) + jj
      for (int g = tid; g < n_ct * KK * 64; g += kNT) {
        const int l = g & 63, kk = (g >> 6) % KK, ct = g / (64 * KK);
        const int c = ct * 32 + (l & 31);
        bf16x8 av;
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          const int pp = 16 * kk + 8 * (l >> 5) + jj;
          float val = 0.f;
          if (pp < D) {
            val = (pp < d && c < k) ? C[(int64_t)pp * k + c] : 0.f;
          } else if (pp < D + 3) {
            if (c < k) {
              const float h = hn[c];
              const float h1 = (float)(__bf16)h;
              const float r1 = h - h1;
              const float h2 = (float)(__bf16)r1;
              const float h3 = (float)(__bf16)(r1 - h2);
              val = pp == D ? h1 : (pp == D + 1 ? h2 : h3);
            } else {
              val = pp == D ? 3.0e38f : 0.f;   // padding centroid: never the best
            }
          }
          av[jj] = (__bf16)val;
        }
        Al[g] = av;
      }
      for (int o = kWave / 2; o > 0; o >>= 1) {
        hmax = fmaxf(hmax, __shfl_xor(hmax, o, kWave));
        hbad |= __shfl_xor(hbad, o, kWave);
      }
      if (lane == 0) {
        red[wave] = hmax;
        bad_hn[wave] = hbad;
      }
      __syncthreads();
      float hnmax = 0.f;
      int any_bad = 0;
#pragma unroll
      for (int i = 0; i < kW; ++i) {
        hnmax = fmaxf(hnmax, red[i]);
        any_bad |= bad_hn[i];
      }
      const float cmax = sqrtf(2.f * hnmax);
      const bool finite_c = !any_bad;
      const int64_t n = F.n;
      const int64_t end = min((int64_t)item.first + item.count, n);
      bf16x8 bh;
#pragma unroll
      for (int jj = 0; jj < 8; ++jj)
        bh[jj] = (__bf16)((half == 0 && jj < 3) ? -1.f : 0.f);
      for (int64_t s0 = item.first; s0 < end; s0 += kW * 32) {  // block-uniform
        const int64_t p = s0 + wave * 32 + col;
        // lane features 16 (i / 8) + 8 half + i % 8
        float xr[XH];
#pragma unroll
        for (int i = 0; i < XH; ++i) {
          const int f = 16 * (i / 8) + 8 * half + (i % 8);
          xr[i] = (f < d && p < end) ? gptr(F.X)[(int64_t)f * n + p] : 0.f;
        }
        float ss = 0.f;
#pragma unroll
        for (int i = 0; i < XH; ++i) ss = fmaf(xr[i], xr[i], ss);
        ss += __shfl_xor(ss, 32, kWave);
        const float xnorm = sqrtf(ss) * 1.0001f;
        bf16x8 bhi[KD];
#pragma unroll
        for (int kd = 0; kd < KD; ++kd)
#pragma unroll
          for (int jj = 0; jj < 8; ++jj) bhi[kd][jj] = (__bf16)xr[8 * kd + jj];
        float b1 = -__builtin_huge_valf(), b2 = -__builtin_huge_valf();
        int bt = 0;
        for (int ct = 0; ct < n_ct; ++ct) {
          const bf16x8 *At = Al + ct * KK * 64 + lane;
          f32x16 acc = {};
#pragma unroll
          for (int kk = 0; kk < KK; ++kk)
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(At[kk * 64],
                                                          kk < KD ? bhi[kk < KD ? kk : 0] : bh,
                                                          acc, 0, 0, 0);
          // rows m = 8 (r / 4) + 4 half + r % 4 of the tile, r packed into
          // the 4 low bits; top-2 of {b1, b2, u, v}: b1' = max3(b1, u, v),
          // b2' = max(b2, med3(b1, u, v)); the best's tile is the last tile
          // in which b1 rose
          const float b1_in = b1;
#pragma unroll
          for (int r = 0; r < 16; r += 2) {
            const float uu =
                __uint_as_float((__float_as_uint(acc[r]) & ~kRowMask) | (unsigned)r);
            const float vv = __uint_as_float((__float_as_uint(acc[r + 1]) & ~kRowMask) |
                                             (unsigned)(r + 1));
            const float mm = __builtin_amdgcn_fmed3f(b1, uu, vv);
            b1 = max3f(b1, uu, vv);
            b2 = __builtin_amdgcn_fmed3f(b2, mm, pinf);
          }
          bt = (b1 != b1_in) ? ct : bt;
        }
        const int rb = (int)(__float_as_uint(b1) & kRowMask);
        int lab = bt * 32 + 8 * (rb >> 2) + 4 * half + (rb & 3);
        {
          const float o1 = __shfl_xor(b1, 32, kWave);
          const float o2 = __shfl_xor(b2, 32, kWave);
          const int ol = __shfl_xor(lab, 32, kWave);
          const float n2 = fmaxf(fminf(b1, o1), fmaxf(b2, o2));
          lab = (o1 > b1) ? ol : lab;
          b1 = fmaxf(b1, o1);
          b2 = n2;
        }
        const float e = xnorm * cmax * alpha + hnmax * beta +
                        pack_err * (fabsf(b1) + fabsf(b2));
        const bool amb = !(b1 - b2 > 2.f * e) || !(xnorm * cmax < 1e36f) || !finite_c ||
                         lab >= k;
        const bool valid = half == 0 && p < end;
        if (valid && !amb) F.labels[p] = lab;
        const bool enq = valid && amb;
        const unsigned long long mask = __ballot(enq);
        if (mask != 0ull) {
          const int cnt = __popcll(mask);
          const int leader = __ffsll((long long)mask) - 1;
          int b0 = 0;
          if (lane == leader) b0 = atomicAdd(&n_list, cnt);
          b0 = __shfl(b0, leader, kWave);
          if (enq) list[b0 + __popcll(mask & ((1ull << lane) - 1ull))] = (int)p;
        }
      }
      __syncthreads();
      // FP32 re-score of the listed samples: bt_assign_krn's chain
      const int nl = n_list;
      for (int j0 = wave * 32; j0 < nl; j0 += kW * 32) {
        const int jl = j0 + col;
        const bool ok = jl < nl;
        const int64_t p = ok ? (int64_t)list[jl] : (int64_t)item.first;
        float xr[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const int f = 2 * ks + half;
          xr[ks] = (f < d && ok) ? gptr(F.X)[(int64_t)f * n + p] : 0.f;
        }
        float bestv = __builtin_huge_valf();
        int besti = 0;
        for (int t = 0; t < n_ct; ++t) {
          const float *At = Cl + t * TF;
          const float *Ht = At + 2 * KS * 32;
          f32x16 acc;
#pragma unroll
          for (int gg = 0; gg < 4; ++gg) {
            const f32x4 h4 = *reinterpret_cast<const f32x4 *>(Ht + 8 * gg + 4 * half);
#pragma unroll
            for (int ee = 0; ee < 4; ++ee) acc[4 * gg + ee] = h4[ee];
          }
#pragma unroll
          for (int k4 = 0; k4 < KS / 4; ++k4) {
            const f32x4 av =
                *reinterpret_cast<const f32x4 *>(At + (k4 * 2 + half) * 128 + col * 4);
#pragma unroll
            for (int jj = 0; jj < 4; ++jj)
              acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[jj], xr[4 * k4 + jj], acc, 0, 0, 0);
          }
#pragma unroll
          for (int g = 0; g < 4; ++g) {
            const int m0 = 8 * g + 4 * half;
#pragma unroll
            for (int ee = 0; ee < 4; ++ee) {
              const float v = -acc[4 * g + ee];
              const bool upd = v < bestv;
              bestv = upd ? v : bestv;
              besti = upd ? (t * 32 + m0 + ee) : besti;
            }
          }
        }
        const float ov = __shfl_xor(bestv, 32, kWave);
        const int oi = __shfl_xor(besti, 32, kWave);
        const bool take = (ov < bestv) || (ov == bestv && oi < besti);
        besti = take ? oi : besti;
        if (half == 0 && ok) F.labels[p] = besti;
      }
    }
  }
}

// ---- update -------------------------------------------------------------------
__global__ void __launch_bounds__(kFB)
    bt_zero_krn(const BFit *__restrict__ fits) {
  const int f = blockIdx.x;
  if (!fits[f].st[0].active) return;
  const BFit F = fits[f];
  const int64_t m = (int64_t)(F.d + 1) * F.k;
  for (int64_t i = threadIdx.x; i < m; i += kFB) F.sums[i] = 0ull;
  for (int c = threadIdx.x; c <= F.k; c += kFB) {
    F.totals[c] = 0;
    F.cursor[c] = 0;
  }
}

__global__ void __launch_bounds__(256)
    bt_hist_krn(const BFit *__restrict__ fits, const CItem *__restrict__ items) {
  __shared__ int h[kMaxK];
  const CItem item = items[blockIdx.x];
  if (!fits[item.fit].st[0].active) return;
  const BFit F = fits[item.fit];
  for (int c = threadIdx.x; c < F.k; c += 256) h[c] = 0;
  int lab[kPerThread];  // all loaded before the first is used (count <= kChunk)
#pragma unroll
  for (int u = 0; u < kPerThread; ++u) {
    const int j = threadIdx.x + u * 256;
    lab[u] = (j < item.count) ? F.labels[item.first + j] : -1;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kPerThread; ++u)
    if (lab[u] >= 0) atomicAdd(&h[lab[u]], 1);
  __syncthreads();
  for (int c = threadIdx.x; c < F.k; c += 256)
    if (h[c] != 0) atomicAdd(&F.totals[c], h[c]);
}

// per fit: start[c] (exclusive scan of totals, start[k] = n) and the first
// segment-sum item of each cluster (items[k] = the fit's item count)
__global__ void __launch_bounds__(kFB)
    bt_cluster_scan_krn(const BFit *__restrict__ fits) {
  __shared__ int wa[kFB / kWave], wb[kFB / kWave];
  const int f = blockIdx.x;
  if (!fits[f].st[0].active) {
    if (threadIdx.x == 0) fits[f].st[0].seg_items = 0;
    return;
  }
  const BFit F = fits[f];
  const int c = threadIdx.x;  // k <= 256
  const int lane = c & (kWave - 1), wave = c / kWave;
  const int va = (c < F.k) ? F.totals[c] : 0;
  const int vb = (va + kMembers - 1) / kMembers;
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
    wa[wave] = ia;
    wb[wave] = ib;
  }
  __syncthreads();
  int pa = 0, pb = 0, ta = 0, tb = 0;
  for (int w2 = 0; w2 < kFB / kWave; ++w2) {
    if (w2 < wave) {
      pa += wa[w2];
      pb += wb[w2];
    }
    ta += wa[w2];
    tb += wb[w2];
  }
  if (c < F.k) {
    F.start[c] = pa + ia - va;
    F.citems[c] = pb + ib - vb;
  }
  if (c == 0) {
    F.start[F.k] = ta;
    F.citems[F.k] = tb;
    fits[f].st[0].seg_items = tb;
  }
}

// exclusive scan of the fits' segment-item counts (one block)
__global__ void __launch_bounds__(1024)
    bt_fit_scan_krn(int n_fits, const BFit *__restrict__ fits, int *__restrict__ total) {
  __shared__ int ws[1024 / kWave];
  __shared__ int carry;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int f0 = 0; f0 < n_fits; f0 += 1024) {
    const int f = f0 + tid;
    const int v = (f < n_fits) ? fits[f].st[0].seg_items : 0;
    int incl = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const int t = __shfl_up(incl, o, kWave);
      if (lane >= o) incl += t;
    }
    if (lane == kWave - 1) ws[wave] = incl;
    __syncthreads();
    int p = 0, tot = 0;
    for (int w2 = 0; w2 < 1024 / kWave; ++w2) {
      if (w2 < wave) p += ws[w2];
      tot += ws[w2];
    }
    if (f < n_fits) fits[f].st[0].seg_base = carry + p + incl - v;
    __syncthreads();
    if (tid == 0) carry += tot;
    __syncthreads();
  }
  if (tid == 0) total[0] = carry;
}

__global__ void __launch_bounds__(256)
    bt_scatter_krn(const BFit *__restrict__ fits, const CItem *__restrict__ items) {
  __shared__ int h[kMaxK];
  const CItem item = items[blockIdx.x];
  if (!fits[item.fit].st[0].active) return;
  const BFit F = fits[item.fit];
  for (int c = threadIdx.x; c < F.k; c += 256) h[c] = 0;
  // the chunk's labels, all loaded before the first is used (count <= kChunk)
  int lab[kPerThread];
#pragma unroll
  for (int u = 0; u < kPerThread; ++u) {
    const int j = threadIdx.x + u * 256;
    lab[u] = (j < item.count) ? F.labels[item.first + j] : -1;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kPerThread; ++u)
    if (lab[u] >= 0) atomicAdd(&h[lab[u]], 1);
  __syncthreads();
  for (int c = threadIdx.x; c < F.k; c += 256)
    h[c] = (h[c] != 0) ? F.start[c] + atomicAdd(&F.cursor[c], h[c]) : 0;
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kPerThread; ++u)
    if (lab[u] >= 0) {
      const int pos = atomicAdd(&h[lab[u]], 1);
      F.perm[pos] = item.first + threadIdx.x + u * 256;
    }
}

// Segment items of every active fit, flattened: map[seg_base + i] = {fit,
// cluster, first member (position in perm), member count} for item i of the
// fit (items of cluster c are citems[c] .. citems[c + 1] - 1, each up to
// kMembers consecutive members of the cluster).
__global__ void __launch_bounds__(kFB)
    bt_seg_map_krn(const BFit *__restrict__ fits, int4 *__restrict__ map) {
  __shared__ int ci[kMaxK + 1];
  const int f = blockIdx.x;
  if (!fits[f].st[0].active) return;
  const BFit F = fits[f];
  const int k = F.k, base = fits[f].st[0].seg_base, items = fits[f].st[0].seg_items;
  for (int c = threadIdx.x; c <= k; c += kFB) ci[c] = F.citems[c];
  __syncthreads();
  for (int i = threadIdx.x; i < items; i += kFB) {
    int a = 0, b = k - 1;  // the last cluster whose first item is <= i
    while (a < b) {
      const int mid = (a + b + 1) >> 1;
      if (ci[mid] <= i)
        a = mid;
      else
        b = mid - 1;
    }
    const int m0 = F.start[a] + (i - ci[a]) * kMembers;
    map[base + i] = make_int4(f, a, m0, min(kMembers, F.start[a + 1] - m0));
  }
}

// One block = one segment item (up to kMembers members of one cluster of one
// fit).  The d + 1 columns (features, then the weight) are split over
// floor(256 / (d + 1)) row groups of d + 1 threads; each thread sums its
// column over every R-th member, the groups are combined in LDS and one
// atomic per column adds the item's sum.  Every term is an integer, so the
// sums do not depend on the order.
constexpr int kSegThreads = 256;
static_assert(kSegThreads / (kMaxD + 1) >= 1, "a row group must fit a block");
__global__ void __launch_bounds__(kSegThreads)
    bt_segsum_krn(const BFit *__restrict__ fits, const int4 *__restrict__ map,
                  const int *__restrict__ total) {
  __shared__ int sidx[kMembers];
  __shared__ float sw[kMembers];
  __shared__ long long red[kSegThreads];
  const int g = blockIdx.x;
  if (g >= total[0]) return;
  const int t = threadIdx.x;
  const int4 e = map[g];
  const BFit F = fits[e.x];
  const int c = e.y, m0 = e.z, cnt = e.w;
  for (int j = t; j < cnt; j += kSegThreads) {
    const int s = F.perm[m0 + j];
    sidx[j] = s;
    sw[j] = F.w[s];
  }
  __syncthreads();
  const auto *Xr = gptr(F.xr);
  const int d = F.d, D1 = d + 1, R = kSegThreads / D1;
  long long acc = 0;
  if (t < R * D1) {
    const int f = t % D1, r = t / D1;
    const float sc = F.scale[f];
    if (f < d) {
      // sixteen row gathers in flight per thread before the first is used
      constexpr int kU = 16;
      int j = r;
      for (; j + (kU - 1) * R < cnt; j += kU * R) {
        int ix[kU];
        float v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) ix[u] = sidx[j + u * R];
#pragma unroll
        for (int u = 0; u < kU; ++u) v[u] = Xr[(int64_t)ix[u] * d + f];
#pragma unroll
        for (int u = 0; u < kU; ++u) acc += __float2int_rn((v[u] * sw[j + u * R]) * sc);
      }
      for (; j < cnt; j += R)
        acc += __float2int_rn((Xr[(int64_t)sidx[j] * d + f] * sw[j]) * sc);
    } else {
      for (int j = r; j < cnt; j += R) acc += __float2int_rn(sw[j] * sc);
    }
  }
  red[t] = acc;
  __syncthreads();
  if (t < D1) {
    long long sum = 0;
    for (int r = 0; r < R; ++r) sum += red[r * D1 + t];
    if (sum != 0) atomicAdd(&F.sums[(int64_t)t * F.k + c], (u64)sum);
  }
}

// ---- empty clusters and relocation ----------------------------------------------
// Ascending list of the fit's empty clusters (a zero int64 count is exactly a
// zero size after conversion) and n_empty.
__global__ void __launch_bounds__(kFB)
    bt_empty_krn(const BFit *__restrict__ fits) {
  __shared__ int wcount[kFB / kWave];
  const int f = blockIdx.x;
  if (!fits[f].st[0].active) return;
  const BFit F = fits[f];
  const int tid = threadIdx.x;  // k <= 256 = kFB
  const bool empty =
      (tid < F.k) && (F.sums[(int64_t)F.d * F.k + tid] == 0ull);
  const uint64_t mask = __ballot(empty);
  const int lane = tid & (kWave - 1), w = tid / kWave;
  const int before_in_wave = __popcll(mask & ((1ull << lane) - 1ull));
  if (lane == 0) wcount[w] = __popcll(mask);
  __syncthreads();
  int before = 0, tot = 0;
  for (int i = 0; i < kFB / kWave; ++i) {
    before += (i < w) ? wcount[i] : 0;
    tot += wcount[i];
  }
  if (empty) F.list[before + before_in_wave] = tid;
  if (tid == 0) fits[f].st[0].n_empty = tot;
}

// squared distance (unit weights) of every sample to its current centroid,
// for fits with an empty cluster (the single-fit inertia arithmetic)
__global__ void __launch_bounds__(256)
    bt_dist_krn(const BFit *__restrict__ fits, const CItem *__restrict__ items) {
  const CItem item = items[blockIdx.x];
  if (!fits[item.fit].st[0].active || fits[item.fit].st[0].n_empty == 0) return;
  const BFit F = fits[item.fit];
  const float *C = cbuf(F, fits[item.fit].st[0].cur);
  for (int j = threadIdx.x; j < item.count; j += 256) {
    const int64_t s = item.first + j;
    const int c = F.labels[s];
    float acc = 0.f;
    constexpr int kU = 16;  // loads in flight; the fmaf chain runs over fe = 0, 1, ...
    for (int f0 = 0; f0 < F.d; f0 += kU) {
      float xv[kU], cv[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        xv[u] = 0.f;
        cv[u] = 0.f;
        if (f0 + u < F.d) {
          xv[u] = F.X[(int64_t)(f0 + u) * F.n + s];
          cv[u] = C[(int64_t)(f0 + u) * F.k + c];
        }
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        if (f0 + u < F.d) {
          const float diff = xv[u] - cv[u];
          acc = fmaf(diff, diff, acc);
        }
      }
    }
    F.dist[s] = acc;
  }
}

// ---- finalize (one block per fit) ----------------------------------------------
__global__ void __launch_bounds__(kFB)
    bt_finalize_krn(const BFit *__restrict__ fits) {
  __shared__ double red[kFB / kWave];
  __shared__ int hist[256];
  __shared__ unsigned sel_prefix;
  __shared__ int sel_left, n_sel, n_eq_taken;
  __shared__ int wcnt[kFB / kWave];
  __shared__ unsigned skey[kMaxK];
  __shared__ int sidx[kMaxK];
  __shared__ int order[kMaxK];
  const int f = blockIdx.x;
  if (!fits[f].st[0].active) return;
  const BFit F = fits[f];
  const int d = F.d, k = F.k, tid = threadIdx.x;
  const int64_t n = F.n;
  const int cur = fits[f].st[0].cur;
  const float *Cc = cbuf(F, cur);
  float *Cn = cbuf(F, cur ^ 1);
  float *sizes = F.sizes;
  const int *list = F.list;
  const float *sc = F.scale;
  const int dk = d * k;  // <= kMaxD * kMaxK
  // Each loop below issues kU loads per thread before using any of them
  // (one round trip per kU elements, not per element); the arithmetic and
  // its order per element are unchanged.
  constexpr int kU = 8;
  // sums -> centroid sums and sizes (the single-slot partial reduction)
  for (int i0 = tid; i0 < dk + k; i0 += kU * kFB) {
    long long q[kU];
    float sv[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int i = i0 + u * kFB;
      q[u] = 0;
      sv[u] = 1.f;
      if (i < dk + k) {
        q[u] = (long long)F.sums[i];
        sv[u] = sc[i / k];
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int i = i0 + u * kFB;
      if (i < dk + k) {
        const float v = (float)((double)q[u] / (double)sv[u]);
        if (i < dk)
          Cn[i] = v;
        else
          sizes[i - dk] = v;
      }
    }
  }
  __syncthreads();
  const int n_empty = fits[f].st[0].n_empty;
  if (n_empty > 0) {
    // the m farthest samples by (distance desc, index asc): radix-select
    // the m-th largest distance (non-negative floats order as their bits),
    // gather the samples above it and the lowest-index ones equal to it,
    // then rank the m of them
    const float *dist = F.dist;
    const int m = (int)min((int64_t)n_empty, n);
    if (tid == 0) {
      sel_prefix = 0u;
      sel_left = m;
    }
    for (int shift = 24; shift >= 0; shift -= 8) {
      hist[tid] = 0;
      __syncthreads();
      const unsigned pre = sel_prefix;
      const unsigned hi_mask = (shift == 24) ? 0u : (0xffffffffu << (shift + 8));
      for (int64_t s = tid; s < n; s += kFB) {
        const unsigned key = __float_as_uint(dist[s]);
        if ((key & hi_mask) == pre) atomicAdd(&hist[(key >> shift) & 255u], 1);
      }
      __syncthreads();
      if (tid == 0) {
        int left = sel_left, b = 255;
        for (; b > 0; --b) {
          if (hist[b] >= left) break;
          left -= hist[b];
        }
        sel_prefix = pre | ((unsigned)b << shift);
        sel_left = left;  // how many of the keys equal to the prefix are taken
      }
      __syncthreads();
    }
    const unsigned thr = sel_prefix;
    const int need_eq = sel_left;
    if (tid == 0) {
      n_sel = 0;
      n_eq_taken = 0;
    }
    __syncthreads();
    // gather in index order (chunks of kFB samples, ballot ranks)
    const int lane = tid & (kWave - 1), w = tid / kWave;
    for (int64_t c0 = 0; c0 < n; c0 += kFB) {
      const int64_t s = c0 + tid;
      const unsigned key = (s < n) ? __float_as_uint(dist[s]) : 0u;
      const bool gt = (s < n) && key > thr;
      const bool eq = (s < n) && key == thr;
      const uint64_t meq = __ballot(eq);
      if (lane == 0) wcnt[w] = __popcll(meq);
      __syncthreads();
      int before = 0, tot = 0;
      for (int i = 0; i < kFB / kWave; ++i) {
        before += (i < w) ? wcnt[i] : 0;
        tot += wcnt[i];
      }
      const int eq_rank = n_eq_taken + before + __popcll(meq & ((1ull << lane) - 1ull));
      const bool take = gt || (eq && eq_rank < need_eq);
      if (take) {
        const int pos = atomicAdd(&n_sel, 1);
        skey[pos] = key;
        sidx[pos] = (int)s;
      }
      __syncthreads();
      if (tid == 0) n_eq_taken += tot;
      __syncthreads();
    }
    for (int i = tid; i < m; i += kFB) {
      int rank = 0;
      for (int j = 0; j < m; ++j)
        rank += (skey[j] > skey[i]) || (skey[j] == skey[i] && sidx[j] < sidx[i]);
      order[rank] = sidx[i];
    }
    __syncthreads();
    // move w x of the chosen samples (in order) into the empty clusters
    for (int fe = tid; fe < d; fe += kFB) {
      for (int j = 0; j < m; ++j) {
        const int s = order[j];
        const int p = F.labels[s];
        const int e = list[j];
        const float ws = F.w[s];
        const float add = F.X[(int64_t)fe * n + s] * ws;
        Cn[(int64_t)fe * k + p] -= add;
        Cn[(int64_t)fe * k + e] = add;
        if (fe == 0) {
          sizes[p] -= ws;
          sizes[e] = ws;
        }
      }
    }
    __syncthreads();
  }
  // divide, then shifts and their sum in the single-fit order
  for (int i0 = tid; i0 < dk; i0 += kU * kFB) {
    float a[kU], b[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int i = i0 + u * kFB;
      a[u] = 0.f;
      b[u] = 1.f;
      if (i < dk) {
        a[u] = Cn[i];
        b[u] = sizes[i % k];
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int i = i0 + u * kFB;
      if (i < dk) Cn[i] = a[u] / b[u];
    }
  }
  __syncthreads();
  double acc = 0.0;
  for (int c = tid; c < k; c += kFB) {
    float s2 = 0.f;
    for (int j0 = 0; j0 < d; j0 += kU) {
      float a[kU], b[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int j = j0 + u;
        a[u] = 0.f;
        b[u] = 0.f;
        if (j < d) {
          a[u] = Cc[(int64_t)j * k + c];
          b[u] = Cn[(int64_t)j * k + c];
        }
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        if (j0 + u < d) {
          const float diff = a[u] - b[u];
          s2 = fmaf(diff, diff, s2);
        }
      }
    }
    acc += (double)s2;
  }
  const double tot = block_sum256(acc, red);
  if (tid == 0) {
    BState S = fits[f].st[0];
    S.shift = tot;
    S.cur = cur ^ 1;
    S.it += 1;
    const float shift_t = (float)tot;
    S.active = (S.it < F.max_iter) && (shift_t > F.tol);
    fits[f].st[0] = S;
  }
}

// ---- epilogue --------------------------------------------------------------------
// per-sample weighted inertia of the final labels
__global__ void __launch_bounds__(256)
    bt_inertia_krn(const BFit *__restrict__ fits, const CItem *__restrict__ items) {
  const CItem item = items[blockIdx.x];
  const BFit F = fits[item.fit];
  const float *C = cbuf(F, fits[item.fit].st[0].cur);
  for (int j = threadIdx.x; j < item.count; j += 256) {
    const int64_t s = item.first + j;
    const int c = F.labels[s];
    float acc = 0.f;
    constexpr int kU = 16;  // loads in flight; the fmaf chain runs over fe = 0, 1, ...
    for (int f0 = 0; f0 < F.d; f0 += kU) {
      float xv[kU], cv[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        xv[u] = 0.f;
        cv[u] = 0.f;
        if (f0 + u < F.d) {
          xv[u] = F.X[(int64_t)(f0 + u) * F.n + s];
          cv[u] = C[(int64_t)(f0 + u) * F.k + c];
        }
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        if (f0 + u < F.d) {
          const float diff = xv[u] - cv[u];
          acc = fmaf(diff, diff, acc);
        }
      }
    }
    F.dist[s] = acc * F.w[s];
  }
}

// stage 1 of the single-fit two-stage sum: block b of fit f sums the
// grid-stride slice i = b*256 + t, step nb_f * 256
__global__ void __launch_bounds__(256)
    bt_sum1_krn(const BFit *__restrict__ fits, const int2 *__restrict__ map,
                const int *__restrict__ nb, double *__restrict__ partials) {
  __shared__ double lds[256 / kWave];
  const int2 fb = map[blockIdx.x];  // (fit, block within the fit)
  const BFit F = fits[fb.x];
  const int nbf = nb[fb.x];
  double acc = 0.0;
  for (int64_t i = (int64_t)fb.y * 256 + threadIdx.x; i < F.n;
       i += (int64_t)nbf * 256)
    acc += (double)F.dist[i];
  const double t = block_sum256(acc, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ void __launch_bounds__(256)
    bt_sum2_krn(const BFit *__restrict__ fits, const int *__restrict__ nb,
                const int *__restrict__ first, const double *__restrict__ partials) {
  __shared__ double lds[256 / kWave];
  const int f = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb[f]; i += 256) acc += partials[first[f] + i];
  const double t = block_sum256(acc, lds);
  if (threadIdx.x == 0) fits[f].st[0].inertia = (double)(float)t;
}

__global__ void __launch_bounds__(256)
    bt_copy_result_krn(const BFit *__restrict__ fits) {
  const int f = blockIdx.x;
  const BFit F = fits[f];
  if (fits[f].st[0].cur == 1) return;  // the result buffer already holds it
  const int64_t m = (int64_t)F.d * F.k;
  for (int64_t i = threadIdx.x; i < m; i += 256) F.C[1][i] = F.C[0][i];
}

// ---- set-up of many new fits at once ---------------------------------------
// (the batched call adds all its fits first; these replace the per-fit
// transpose, maxima and scales launches with one launch each)

// state (active while max_iter > 0) and cleared maxima of each new fit
__global__ void __launch_bounds__(256)
    bt_prep_state_krn(int m, const BFit *__restrict__ fits) {
  const int f = blockIdx.x * 256 + threadIdx.x;
  if (f >= m) return;
  const BFit F = fits[f];
  BState S{};
  S.active = F.max_iter > 0 ? 1 : 0;
  F.st[0] = S;
  for (int j = 0; j < range_maxbits_count(F.d); ++j) F.maxbits[j] = 0u;
  for (int j = 0; j < range_estat_count(F.d); ++j) F.estat[j] = 0ull;
}

// X_r = transpose of X_t for one 32 x 32 tile of one fit (transpose_krn's
// tile, per item)
__global__ void __launch_bounds__(256)
    bt_transpose_krn(const BFit *__restrict__ fits, const TItem *__restrict__ items) {
  __shared__ float tile[32][33];
  const TItem it = items[blockIdx.x];
  const BFit F = fits[it.fit];
  const int64_t n = F.n;
  const int d = F.d;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    const int f = it.f0 + r;
    const int64_t s = (int64_t)it.s0 + tx;
    tile[r][tx] = (f < d && s < n) ? F.X[(int64_t)f * n + s] : 0.f;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int64_t s = (int64_t)it.s0 + r;
    const int f = it.f0 + tx;
    if (s < n && f < d) F.xr[s * d + f] = tile[tx][r];
  }
}

// The range statistics of fixed_point.hpp over the item's samples (f == d:
// the weights), as absmax_bits_krn gathers them per fit (integer maxima and
// sums: the same values in any order)
__global__ void __launch_bounds__(256)
    bt_absmax_krn(const BFit *__restrict__ fits, const CItem *__restrict__ items) {
  const CItem it = items[blockIdx.x];
  const BFit F = fits[it.fit];
  for (int f = 0; f <= F.d; ++f) {
    const float *row = (f < F.d) ? F.X + (int64_t)f * F.n : F.w;
    unsigned m = 0u, esum = 0u, ecnt = 0u, wmin_inv = 0u;
    for (int i = threadIdx.x; i < it.count; i += 256) {
      const unsigned v = __float_as_uint(row[it.first + i]);
      range_accumulate(v, m, esum, ecnt);
      const unsigned b = v & 0x7fffffffu;
      if (f == F.d && b != 0u) wmin_inv = max(wmin_inv, ~b);
    }
    u64 es = esum, ec = ecnt;
    for (int o = kWave / 2; o > 0; o >>= 1) {
      m = max(m, (unsigned)__shfl_xor((int)m, o, kWave));
      wmin_inv = max(wmin_inv, (unsigned)__shfl_xor((int)wmin_inv, o, kWave));
      es += __shfl_xor(es, o, kWave);
      ec += __shfl_xor(ec, o, kWave);
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
      atomicMax(&F.maxbits[f], m);
      if (f == F.d && wmin_inv != 0u) atomicMax(&F.maxbits[F.d + 1], wmin_inv);
      if (ec != 0) {
        atomicAdd(&F.estat[2 * f], es);
        atomicAdd(&F.estat[2 * f + 1], ec);
      }
    }
  }
}

// each new fit's fixed-point scales and non-finite flag (fixed_point.hpp)
__global__ void __launch_bounds__(256)
    bt_scales_krn(int m, const BFit *__restrict__ fits) {
  const int f = blockIdx.x * 256 + threadIdx.x;
  if (f >= m) return;
  const BFit F = fits[f];
  fixed_point_scales_from_bits(F.n, F.d, F.maxbits, F.estat, F.scale, F.nonfinite);
}

// out[i] = *flags[i]: the new fits' non-finite flags read back with one copy
__global__ void __launch_bounds__(256)
    bt_gather_flags_krn(int m, const int *const *__restrict__ flags, int *out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < m) out[i] = *flags[i];
}

// out[f] = fit f's state: the host reads every fit's state with one copy
__global__ void __launch_bounds__(256)
    bt_gather_state_krn(int nf, const BFit *__restrict__ fits, BState *out) {
  const int f = blockIdx.x * 256 + threadIdx.x;
  if (f < nf) out[f] = fits[f].st[0];
}

// ---- host ---------------------------------------------------------------------
inline int ks_bucket(int d) { return d <= 16 ? 8 : (d <= 32 ? 16 : 32); }

size_t assign_lds(int KS) { return (size_t)(kMaxK / 32) * (2 * KS * 32 + 32) * 4; }

// Assignment launch shapes the batched call tunes between (results never
// depend on them): threads per block = assign_threads(KS) (shape 0) or half
// of it (shape 1).
template <int KS, int NT>
int assign_grid_nt(bool mixed) {
  static int grid[2] = {0, 0};
  int &g = grid[mixed ? 1 : 0];
  if (g == 0) {
    const void *fn = mixed ? reinterpret_cast<const void *>(&bt_assign_mixed_krn<KS, NT>)
                           : reinterpret_cast<const void *>(&bt_assign_krn<KS, NT>);
    const size_t lds = mixed ? mixed_assign_lds<KS>() : assign_lds(KS);
    KMD_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    int per_cu = 0;
    KMD_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NT, lds));
    g = device_cu_count() * std::max(1, per_cu);
  }
  return g;
}

template <int KS>
int assign_grid() {
  return assign_grid_nt<KS, assign_threads(KS)>(false);
}

template <int KS>
int assign_launch_grid(int n_items) {
  return std::max(1, std::min(assign_grid<KS>(), n_items));
}

template <int KS, int NT>
void launch_assign_nt(const BFit *fits, const AItem *items, int n_items, int all,
                      hipStream_t s, bool mixed) {
  const int grid = std::max(1, std::min(assign_grid_nt<KS, NT>(mixed), n_items));
  if (mixed)
    hipLaunchKernelGGL((bt_assign_mixed_krn<KS, NT>), dim3(grid), dim3(NT),
                       mixed_assign_lds<KS>(), s, fits, items, n_items, all,
                       __builtin_huge_valf());
  else
    hipLaunchKernelGGL((bt_assign_krn<KS, NT>), dim3(grid), dim3(NT), assign_lds(KS), s,
                       fits, items, n_items, all);
  KMD_LAUNCH_CHECK();
}

// mixed: the BF16 cross term with the certified FP32 re-score (the same
// labels as the FP32 kernel) for the d <= 32 buckets; the d <= 64 bucket
// runs FP32 either way (its BF16 kernel measured slower: 430 against
// 385 us per launch on the 512-fit workload).  shape: 0 or 1
// (assign_grid_nt)
template <int KS>
void launch_assign_bucket(const BFit *fits, const AItem *items, int n_items,
                          int all, hipStream_t s, bool mixed, int shape) {
  if (n_items <= 0) return;
  mixed = mixed && KS <= 16;
  if (shape == 1)
    launch_assign_nt<KS, assign_threads(KS) / 2>(fits, items, n_items, all, s, mixed);
  else
    launch_assign_nt<KS, assign_threads(KS)>(fits, items, n_items, all, s, mixed);
}

inline int assign_bucket(int d) { return ks_bucket(d) == 8 ? 0 : (ks_bucket(d) == 16 ? 1 : 2); }

// Append fit f's assignment items to its bucket's list.
void append_assign_items(const BFit &F, int f, std::vector<AItem> *ai) {
  const int64_t macs = (int64_t)((F.k + 31) / 32 * 32) * (2 * ks_bucket(F.d));
  int64_t len = kItemMacs / macs;
  len = std::max<int64_t>(kItemMin, std::min<int64_t>(kItemMax, len)) / kItemMin * kItemMin;
  for (int64_t p = 0; p < F.n; p += len)
    ai[assign_bucket(F.d)].push_back(
        AItem{f, (int)p, (int)std::min<int64_t>(len, F.n - p)});
}

// Order one bucket's items for the grid-stride loop of `grid` blocks:
// heaviest first (work = samples x padded centroids; ties by fit, first),
// then each round of `grid` items reversed every other round, so every
// block's total work comes out about equal.
void order_assign_items(const std::vector<BFit> &fits, int grid, std::vector<AItem> &v) {
  auto work = [&](const AItem &a) { return (int64_t)a.count * ((fits[a.fit].k + 31) / 32); };
  std::stable_sort(v.begin(), v.end(),
                   [&](const AItem &a, const AItem &b) { return work(a) > work(b); });
  if (grid <= 1) return;
  for (size_t r0 = (size_t)grid; r0 < v.size(); r0 += 2 * (size_t)grid)
    std::reverse(v.begin() + r0, v.begin() + std::min(v.size(), r0 + grid));
}

void order_all_assign_items(const std::vector<BFit> &fits, std::vector<AItem> *ai) {
  order_assign_items(fits, assign_launch_grid<8>((int)ai[0].size()), ai[0]);
  order_assign_items(fits, assign_launch_grid<16>((int)ai[1].size()), ai[1]);
  order_assign_items(fits, assign_launch_grid<32>((int)ai[2].size()), ai[2]);
}

__global__ void bt_set_state_krn(BState *st, int active) {
  BState S{};
  S.active = active;
  st[0] = S;
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Byte layout of one fit's workspace arena.
struct Arena {
  size_t st, xr, perm, dist, sums, scale, hn, sizes, list, totals, cursor, start,
      citems, nonfinite, maxbits, estat, total;
  Arena(int64_t n, int64_t d, int64_t k) {
    size_t o = 0;
    auto take = [&](size_t b) {
      const size_t at = o;
      o += align256(b);
      return at;
    };
    st = take(sizeof(BState));
    xr = take(sizeof(float) * (size_t)n * d);
    perm = take(sizeof(int) * (size_t)n);
    dist = take(sizeof(float) * (size_t)n);
    sums = take(sizeof(u64) * (size_t)(d + 1) * k);
    scale = take(sizeof(float) * (size_t)(d + 1));
    hn = take(sizeof(float) * (size_t)(k + 1));
    sizes = take(sizeof(float) * (size_t)(k + 1));
    list = take(sizeof(int) * (size_t)(k + 1));
    totals = take(sizeof(int) * (size_t)(k + 1));
    cursor = take(sizeof(int) * (size_t)(k + 1));
    start = take(sizeof(int) * (size_t)(k + 1));
    citems = take(sizeof(int) * (size_t)(k + 1));
    nonfinite = take(sizeof(int));
    maxbits = take(sizeof(unsigned) * (size_t)range_maxbits_count(d));
    estat = take(sizeof(u64) * (size_t)range_estat_count(d));
    total = o;
  }
};

}  // namespace

bool batched_fit_supported(int64_t n, int64_t d, int64_t k) {
  return n >= 1 && n < (int64_t(1) << 30) && d >= 1 && d <= kMaxD && k >= 1 &&
         k <= kMaxK && k <= n;
}

int64_t batch_fit_footprint(int64_t n, int64_t d, int64_t k) {
  return (int64_t)Arena(n, d, k).total;
}

// A running set of fits: fits join with add(), step() advances every active
// fit by whole iterations (one launch per step of an iteration for all of
// them; replayed from a captured hipGraph while the set is unchanged) and
// runs the epilogue of fits that stopped; finished fits leave the set (and
// free their workspace) at the next step after take_finished().
struct BatchEngine {
  struct Slot;
  explicit BatchEngine(hipStream_t user_stream);
  ~BatchEngine();
  BatchEngine(const BatchEngine &) = delete;
  BatchEngine &operator=(const BatchEngine &) = delete;
  // defer_prep: the fit's set-up (transpose, maxima, scales) waits for
  // prep_pending(), which does it for all deferred fits in a few launches
  // arena: when given, the fit's workspace (arena_bytes(f) bytes, 256-byte
  // aligned) instead of an allocation of its own
  int64_t add(const BatchedFit &f, bool check_now, bool defer_prep = false,
              char *arena = nullptr);
  static size_t arena_bytes(const BatchedFit &f);
  void prep_pending();
  // defer_epilogue: fits that stop are set aside and get their epilogue
  // together at flush_stopped() (the one-call batched path)
  bool defer_epilogue_ = false;
  std::vector<Slot *> stopped_;
  void flush_stopped();
  int64_t step(int iterations, bool use_graph);
  std::vector<int64_t> take_finished();
  int64_t resident_bytes() const;
  void rebuild();
  // ev_a0 / ev_a1 (may be null): recorded around the assignment launches
  void iteration(hipStream_t q, hipEvent_t ev_a0 = nullptr, hipEvent_t ev_a1 = nullptr);
  void epilogue(const std::vector<Slot *> &done);
  void check_inputs();
  void drop_graph();
  // fits whose data range needs floating-point sums (fixed_point.hpp flag 2):
  // set aside for the single-fit driver when route_wide_ is set, else add()
  // / step() raise ValueError
  bool route_wide_ = false;
  std::vector<BatchedFit> wide_;
  // assignment with the BF16 cross term + certified FP32 re-score
  bool mixed_ = false;
  // launch shape of the assignment kernels (launch_assign_bucket); with
  // tune_shape_ set, the first two iterations of step() run shapes 0 and 1
  // (ungraphed, their assignment launches timed) and the faster is kept
  int assign_shape_ = 0;
  bool tune_shape_ = false;
  float shape_ms_[2] = {0.f, 0.f};

  hipStream_t user_ = nullptr, s_ = nullptr;
  hipEvent_t ev_ = nullptr;
  std::vector<std::unique_ptr<Slot>> slots_;
  std::vector<Slot *> unchecked_, pending_prep_;
  std::vector<BFit> host_fits_;
  std::vector<AItem> ai_[3];
  std::vector<CItem> ci_;
  std::unique_ptr<TempBuffer> fits_d_, ai_d_[3], ci_d_, seg_total_, seg_map_, st_d_;
  // every fit's state in one copy: gathered on the device, read back once
  void read_states(const BFit *dfits, int nf, std::vector<BState> &out);
  int64_t seg_bound_ = 1, next_id_ = 0, launches_ = 0;
  bool dirty_ = true;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  std::vector<int64_t> finished_;
};

struct BatchEngine::Slot {
  BatchedFit desc;
  int64_t id;
  TempBuffer arena{0, nullptr};
  BFit dev{};
  bool done = false;
  bool stopped = false;  // stopped, epilogue deferred
};

BatchEngine::BatchEngine(hipStream_t user_stream) : user_(user_stream) {
  KMD_HIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
  KMD_HIP_CHECK(hipEventCreateWithFlags(&ev_, hipEventDisableTiming));
  (void)assign_grid<8>();
  (void)assign_grid<16>();
  (void)assign_grid<32>();

}

BatchEngine::~BatchEngine() {
  (void)hipStreamSynchronize(s_);
  drop_graph();
  // every stream-ordered buffer is freed before its stream goes away
  unchecked_.clear();
  slots_.clear();
  fits_d_.reset();
  for (auto &b : ai_d_) b.reset();
  ci_d_.reset();
  seg_total_.reset();
  seg_map_.reset();
  st_d_.reset();
  (void)hipStreamSynchronize(s_);
  (void)hipEventDestroy(ev_);
  (void)hipStreamDestroy(s_);
}

void BatchEngine::drop_graph() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
}

// check_now: verify the fit's inputs are finite before returning (one host
// synchronisation); otherwise that happens at the next step for all new fits.
size_t BatchEngine::arena_bytes(const BatchedFit &f) {
  return Arena(f.n, f.d, f.k).total;
}

int64_t BatchEngine::add(const BatchedFit &f, bool check_now, bool defer_prep,
                         char *arena) {
  require(batched_fit_supported(f.n, f.d, f.k), "batched fit: unsupported shape");
  auto slot = std::make_unique<Slot>();
  slot->desc = f;
  slot->id = next_id_++;
  const Arena A(f.n, f.d, f.k);
  if (!defer_prep) {
    // the caller's arrays are ready once its stream reaches this point
    KMD_HIP_CHECK(hipEventRecord(ev_, user_));
    KMD_HIP_CHECK(hipStreamWaitEvent(s_, ev_, 0));
  }
  char *b = arena;
  if (b == nullptr) {
    slot->arena = TempBuffer(A.total, s_);
    b = slot->arena.as<char>();
  }
  BFit &F = slot->dev;
  F.n = f.n;
  F.d = (int)f.d;
  F.k = (int)f.k;
  F.X = f.X;
  F.w = f.w;
  F.C[0] = f.init;
  F.C[1] = f.res;
  F.labels = f.labels;
  F.max_iter = f.max_iter;
  F.tol = (float)f.tol;
  F.st = reinterpret_cast<BState *>(b + A.st);
  F.xr = reinterpret_cast<float *>(b + A.xr);
  F.perm = reinterpret_cast<int *>(b + A.perm);
  F.dist = reinterpret_cast<float *>(b + A.dist);
  F.sums = reinterpret_cast<u64 *>(b + A.sums);
  F.scale = reinterpret_cast<float *>(b + A.scale);
  F.hn = reinterpret_cast<float *>(b + A.hn);
  F.sizes = reinterpret_cast<float *>(b + A.sizes);
  F.list = reinterpret_cast<int *>(b + A.list);
  F.totals = reinterpret_cast<int *>(b + A.totals);
  F.cursor = reinterpret_cast<int *>(b + A.cursor);
  F.start = reinterpret_cast<int *>(b + A.start);
  F.citems = reinterpret_cast<int *>(b + A.citems);
  F.nonfinite = reinterpret_cast<int *>(b + A.nonfinite);
  F.maxbits = reinterpret_cast<unsigned *>(b + A.maxbits);
  F.estat = reinterpret_cast<u64 *>(b + A.estat);
  if (defer_prep) {
    pending_prep_.push_back(slot.get());
  } else {
    hipLaunchKernelGGL(bt_set_state_krn, dim3(1), dim3(1), 0, s_, F.st,
                       f.max_iter > 0 ? 1 : 0);
    KMD_LAUNCH_CHECK();
    // sample-major copy and the single-fit driver's fixed-point scales
    launch_transpose(f.n, f.d, f.X, F.xr, s_);
    launch_fixed_point_scales(f.n, f.d, f.X, f.w, F.scale, F.nonfinite, s_);
  }
  unchecked_.push_back(slot.get());
  slots_.push_back(std::move(slot));
  dirty_ = true;
  const int64_t id = slots_.back()->id;
  if (check_now) check_inputs();
  return id;
}

void BatchEngine::prep_pending() {
  if (pending_prep_.empty()) return;
  // the caller's arrays are ready once its stream reaches this point
  KMD_HIP_CHECK(hipEventRecord(ev_, user_));
  KMD_HIP_CHECK(hipStreamWaitEvent(s_, ev_, 0));
  const int m = (int)pending_prep_.size();
  std::vector<BFit> fits(m);
  std::vector<TItem> ti;
  std::vector<CItem> ci;
  for (int f = 0; f < m; ++f) {
    const BFit &F = pending_prep_[f]->dev;
    fits[f] = F;
    for (int64_t s0 = 0; s0 < F.n; s0 += 32)
      for (int f0 = 0; f0 < F.d; f0 += 32) ti.push_back(TItem{f, (int)s0, f0});
    for (int64_t p = 0; p < F.n; p += kChunk)
      ci.push_back(CItem{f, (int)p, (int)std::min<int64_t>(kChunk, F.n - p)});
  }
  TempBuffer dF(sizeof(BFit) * m, s_), dT(sizeof(TItem) * std::max<size_t>(1, ti.size()), s_),
      dC(sizeof(CItem) * std::max<size_t>(1, ci.size()), s_);
  KMD_HIP_CHECK(hipMemcpyAsync(dF.as<void>(), fits.data(), sizeof(BFit) * m,
                               hipMemcpyHostToDevice, s_));
  if (!ti.empty())
    KMD_HIP_CHECK(hipMemcpyAsync(dT.as<void>(), ti.data(), sizeof(TItem) * ti.size(),
                                 hipMemcpyHostToDevice, s_));
  if (!ci.empty())
    KMD_HIP_CHECK(hipMemcpyAsync(dC.as<void>(), ci.data(), sizeof(CItem) * ci.size(),
                                 hipMemcpyHostToDevice, s_));
  const unsigned gm = (unsigned)ceil_div<int64_t>(m, 256);
  hipLaunchKernelGGL(bt_prep_state_krn, dim3(gm), dim3(256), 0, s_, m, dF.as<const BFit>());
  KMD_LAUNCH_CHECK();
  if (!ti.empty()) {
    hipLaunchKernelGGL(bt_transpose_krn, dim3((unsigned)ti.size()), dim3(256), 0, s_,
                       dF.as<const BFit>(), dT.as<const TItem>());
    KMD_LAUNCH_CHECK();
  }
  if (!ci.empty()) {
    hipLaunchKernelGGL(bt_absmax_krn, dim3((unsigned)ci.size()), dim3(256), 0, s_,
                       dF.as<const BFit>(), dC.as<const CItem>());
    KMD_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(bt_scales_krn, dim3(gm), dim3(256), 0, s_, m, dF.as<const BFit>());
  KMD_LAUNCH_CHECK();
  // the host vectors are copied before they go out of scope
  KMD_HIP_CHECK(hipStreamSynchronize(s_));
  pending_prep_.clear();
}

void BatchEngine::check_inputs() {
  prep_pending();
  if (unchecked_.empty()) return;
  const int m = (int)unchecked_.size();
  std::vector<int> bad(m);
  std::vector<const int *> ptrs(m);
  for (int i = 0; i < m; ++i) ptrs[i] = unchecked_[i]->dev.nonfinite;
  {
    TempBuffer dp(sizeof(const int *) * m, s_), dv(sizeof(int) * m, s_);
    KMD_HIP_CHECK(hipMemcpyAsync(dp.as<void>(), ptrs.data(), sizeof(const int *) * m,
                                 hipMemcpyHostToDevice, s_));
    hipLaunchKernelGGL(bt_gather_flags_krn, dim3((unsigned)ceil_div<int64_t>(m, 256)),
                       dim3(256), 0, s_, m, dp.as<const int *const>(), dv.as<int>());
    KMD_LAUNCH_CHECK();
    KMD_HIP_CHECK(hipMemcpyAsync(bad.data(), dv.as<int>(), sizeof(int) * m,
                                 hipMemcpyDeviceToHost, s_));
    KMD_HIP_CHECK(hipStreamSynchronize(s_));
  }
  std::vector<Slot *> failed;
  bool nonfinite = false, wide = false;
  for (size_t i = 0; i < unchecked_.size(); ++i)
    if (bad[i]) {
      failed.push_back(unchecked_[i]);
      nonfinite |= bad[i] == 1;
      wide |= bad[i] == 2;
      if (bad[i] == 2 && route_wide_) wide_.push_back(unchecked_[i]->desc);
    }
  unchecked_.clear();
  if (!failed.empty()) {
    for (Slot *f : failed)
      slots_.erase(std::remove_if(slots_.begin(), slots_.end(),
                                  [&](const std::unique_ptr<Slot> &p) {
                                    return p.get() == f;
                                  }),
                   slots_.end());
    dirty_ = true;
    if (nonfinite) throw ValueError("batched fit: X or sample_weight holds inf/NaN");
    if (wide && !route_wide_)
      throw ValueError("batched fit: the range of X or sample_weight needs "
                       "floating-point sums; use the single-fit driver");
  }
}

void BatchEngine::rebuild() {
  // drop finished fits, then flatten the rest into work lists
  slots_.erase(std::remove_if(slots_.begin(), slots_.end(),
                              [](const std::unique_ptr<Slot> &p) { return p->done; }),
               slots_.end());
  drop_graph();
  host_fits_.clear();
  for (auto &v : ai_) v.clear();
  ci_.clear();
  seg_bound_ = 1;
  for (size_t f = 0; f < slots_.size(); ++f) {
    const BFit &F = slots_[f]->dev;
    host_fits_.push_back(F);
    append_assign_items(F, (int)f, ai_);
    for (int64_t first = 0; first < F.n; first += kChunk)
      ci_.push_back(CItem{(int)f, (int)first, (int)std::min<int64_t>(kChunk, F.n - first)});
    seg_bound_ += F.n / kMembers + F.k;
  }
  order_all_assign_items(host_fits_, ai_);
  auto upload = [&](std::unique_ptr<TempBuffer> &buf, const void *src, size_t bytes) {
    buf = std::make_unique<TempBuffer>(std::max<size_t>(bytes, 64), s_);
    if (bytes) KMD_HIP_CHECK(hipMemcpyAsync(buf->as<void>(), src, bytes,
                                            hipMemcpyHostToDevice, s_));
  };
  upload(fits_d_, host_fits_.data(), sizeof(BFit) * host_fits_.size());
  for (int b = 0; b < 3; ++b) upload(ai_d_[b], ai_[b].data(), sizeof(AItem) * ai_[b].size());
  upload(ci_d_, ci_.data(), sizeof(CItem) * ci_.size());
  if (!seg_total_) seg_total_ = std::make_unique<TempBuffer>(256, s_);
  seg_map_ = std::make_unique<TempBuffer>(sizeof(int4) * (size_t)seg_bound_, s_);
  KMD_HIP_CHECK(hipStreamSynchronize(s_));  // host vectors may change after this
  dirty_ = false;
}

void BatchEngine::iteration(hipStream_t q, hipEvent_t ev_a0, hipEvent_t ev_a1) {
  const BFit *dF = fits_d_->as<const BFit>();
  const int nf = (int)host_fits_.size();
  const int nci = (int)ci_.size();
  if (nf == 0) return;
  hipLaunchKernelGGL(bt_hn_krn, dim3(nf), dim3(kFB), 0, q, dF, 0);
  KMD_LAUNCH_CHECK();
  if (ev_a0 != nullptr) KMD_HIP_CHECK(hipEventRecord(ev_a0, q));
  launch_assign_bucket<8>(dF, ai_d_[0]->as<AItem>(), (int)ai_[0].size(), 0, q, mixed_,
                          assign_shape_);
  launch_assign_bucket<16>(dF, ai_d_[1]->as<AItem>(), (int)ai_[1].size(), 0, q, mixed_,
                           assign_shape_);
  launch_assign_bucket<32>(dF, ai_d_[2]->as<AItem>(), (int)ai_[2].size(), 0, q, mixed_,
                           assign_shape_);
  if (ev_a1 != nullptr) KMD_HIP_CHECK(hipEventRecord(ev_a1, q));
  hipLaunchKernelGGL(bt_zero_krn, dim3(nf), dim3(kFB), 0, q, dF);
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(bt_hist_krn, dim3(nci), dim3(256), 0, q, dF, ci_d_->as<CItem>());
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(bt_cluster_scan_krn, dim3(nf), dim3(kFB), 0, q, dF);
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(bt_fit_scan_krn, dim3(1), dim3(1024), 0, q, nf, dF,
                     seg_total_->as<int>());
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(bt_scatter_krn, dim3(nci), dim3(256), 0, q, dF,
                     ci_d_->as<CItem>());
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(bt_seg_map_krn, dim3(nf), dim3(kFB), 0, q, dF, seg_map_->as<int4>());
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(bt_segsum_krn, dim3((unsigned)seg_bound_), dim3(kSegThreads), 0, q, dF,
                     seg_map_->as<const int4>(), seg_total_->as<int>());
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(bt_empty_krn, dim3(nf), dim3(kFB), 0, q, dF);
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(bt_dist_krn, dim3(nci), dim3(256), 0, q, dF, ci_d_->as<CItem>());
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(bt_finalize_krn, dim3(nf), dim3(kFB), 0, q, dF);
  KMD_LAUNCH_CHECK();
}

// Labels and weighted inertia of the final centroids for `done` fits, the
// result copied into res when it ended in init (the single-fit epilogue).
void BatchEngine::epilogue(const std::vector<Slot *> &done) {
  if (done.empty()) return;
  std::vector<BFit> fits;
  std::vector<AItem> ai[3];
  std::vector<CItem> ci;
  std::vector<int> nb, first;
  std::vector<int2> map;
  for (size_t f = 0; f < done.size(); ++f) {
    const BFit &F = done[f]->dev;
    fits.push_back(F);
    append_assign_items(F, (int)f, ai);
    for (int64_t p = 0; p < F.n; p += kChunk)
      ci.push_back(CItem{(int)f, (int)p, (int)std::min<int64_t>(kChunk, F.n - p)});
    int64_t g = (F.n + 4095) / 4096;
    g = std::max<int64_t>(1, std::min<int64_t>(1024, g));
    nb.push_back((int)g);
    first.push_back((int)map.size());
    for (int b = 0; b < g; ++b) map.push_back(make_int2((int)f, b));
  }
  order_all_assign_items(fits, ai);
  const int nf = (int)fits.size();
  auto up = [&](const void *src, size_t bytes) {
    auto buf = std::make_unique<TempBuffer>(std::max<size_t>(bytes, 64), s_);
    if (bytes) KMD_HIP_CHECK(hipMemcpyAsync(buf->as<void>(), src, bytes,
                                            hipMemcpyHostToDevice, s_));
    return buf;
  };
  auto dF = up(fits.data(), sizeof(BFit) * fits.size());
  std::unique_ptr<TempBuffer> dai[3];
  for (int b = 0; b < 3; ++b) dai[b] = up(ai[b].data(), sizeof(AItem) * ai[b].size());
  auto dci = up(ci.data(), sizeof(CItem) * ci.size());
  auto dmap = up(map.data(), sizeof(int2) * map.size());
  auto dnb = up(nb.data(), sizeof(int) * nb.size());
  auto dfirst = up(first.data(), sizeof(int) * first.size());
  TempBuffer part(sizeof(double) * std::max<size_t>(1, map.size()), s_);
  const BFit *F = dF->as<const BFit>();
  hipLaunchKernelGGL(bt_hn_krn, dim3(nf), dim3(kFB), 0, s_, F, 1);
  KMD_LAUNCH_CHECK();
  launch_assign_bucket<8>(F, dai[0]->as<AItem>(), (int)ai[0].size(), 1, s_, mixed_,
                          assign_shape_);
  launch_assign_bucket<16>(F, dai[1]->as<AItem>(), (int)ai[1].size(), 1, s_, mixed_,
                           assign_shape_);
  launch_assign_bucket<32>(F, dai[2]->as<AItem>(), (int)ai[2].size(), 1, s_, mixed_,
                           assign_shape_);
  hipLaunchKernelGGL(bt_inertia_krn, dim3((unsigned)ci.size()), dim3(256), 0, s_, F,
                     dci->as<CItem>());
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(bt_sum1_krn, dim3((unsigned)map.size()), dim3(256), 0, s_, F,
                     dmap->as<int2>(), dnb->as<int>(), part.as<double>());
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(bt_sum2_krn, dim3(nf), dim3(256), 0, s_, F, dnb->as<int>(),
                     dfirst->as<int>(), part.as<double>());
  KMD_LAUNCH_CHECK();
  hipLaunchKernelGGL(bt_copy_result_krn, dim3(nf), dim3(256), 0, s_, F);
  KMD_LAUNCH_CHECK();
  std::vector<BState> st;
  read_states(F, nf, st);
  for (int f = 0; f < nf; ++f) {
    Slot *sl = done[f];
    if (sl->desc.out_iters) *sl->desc.out_iters = st[f].it;
    if (sl->desc.out_inertia) *sl->desc.out_inertia = st[f].inertia;
    sl->done = true;
    finished_.push_back(sl->id);
  }
}

void BatchEngine::read_states(const BFit *dfits, int nf, std::vector<BState> &out) {
  out.resize(nf);
  if (nf == 0) return;
  const size_t bytes = sizeof(BState) * (size_t)nf;
  if (!st_d_ || st_d_->bytes() < bytes)
    st_d_ = std::make_unique<TempBuffer>(bytes, s_);
  hipLaunchKernelGGL(bt_gather_state_krn, dim3((unsigned)ceil_div<int64_t>(nf, 256)),
                     dim3(256), 0, s_, nf, dfits, st_d_->as<BState>());
  KMD_LAUNCH_CHECK();
  KMD_HIP_CHECK(hipMemcpyAsync(out.data(), st_d_->as<BState>(), bytes,
                               hipMemcpyDeviceToHost, s_));
  KMD_HIP_CHECK(hipStreamSynchronize(s_));
}

int64_t BatchEngine::step(int iterations, bool use_graph) {
  check_inputs();
  if (dirty_) rebuild();
  if (host_fits_.empty()) return 0;
  // the launch-shape trial: one timed iteration per shape (each a real
  // iteration of every active fit; the shapes give the same results)
  int trial = 0;
  if (tune_shape_) {
    hipEvent_t e0, e1;
    KMD_HIP_CHECK(hipEventCreate(&e0));
    KMD_HIP_CHECK(hipEventCreate(&e1));
    for (int sh = 0; sh < 2 && trial < iterations; ++sh, ++trial) {
      assign_shape_ = sh;
      iteration(s_, e0, e1);
      KMD_HIP_CHECK(hipEventSynchronize(e1));
      KMD_HIP_CHECK(hipEventElapsedTime(&shape_ms_[sh], e0, e1));
    }
    KMD_HIP_CHECK(hipEventDestroy(e0));
    KMD_HIP_CHECK(hipEventDestroy(e1));
    if (trial == 2) {
      tune_shape_ = false;
      assign_shape_ = shape_ms_[1] < shape_ms_[0] ? 1 : 0;
      drop_graph();
    } else {
      assign_shape_ = 0;
    }
    iterations -= trial;
    launches_ += trial;
  }
  if (use_graph && !exec_) {
    KMD_HIP_CHECK(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
    iteration(s_);
    KMD_HIP_CHECK(hipStreamEndCapture(s_, &graph_));
    KMD_HIP_CHECK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
  }
  for (int i = 0; i < iterations; ++i) {
    if (use_graph)
      KMD_HIP_CHECK(hipGraphLaunch(exec_, s_));
    else
      iteration(s_);
  }
  launches_ += iterations;
  // read the masks; fits that stopped get their epilogue
  std::vector<BState> st;
  read_states(fits_d_->as<const BFit>(), (int)host_fits_.size(), st);
  std::vector<Slot *> stopping;
  int64_t active = 0;
  for (size_t f = 0; f < slots_.size(); ++f) {
    if (slots_[f]->done || slots_[f]->stopped) continue;
    if (st[f].active)
      ++active;
    else
      stopping.push_back(slots_[f].get());
  }
  if (defer_epilogue_) {
    for (Slot *sl : stopping) {
      sl->stopped = true;
      stopped_.push_back(sl);
    }
  } else {
    epilogue(stopping);
  }
  return active;
}

void BatchEngine::flush_stopped() {
  epilogue(stopped_);
  stopped_.clear();
}

std::vector<int64_t> BatchEngine::take_finished() {
  std::vector<int64_t> out;
  out.swap(finished_);
  if (!out.empty()) dirty_ = true;  // free their workspaces at the next step
  return out;
}

int64_t BatchEngine::resident_bytes() const {
  int64_t b = 0;
  for (const auto &s : slots_)
    if (!s->done) b += (int64_t)s->arena.bytes();
  return b;
}

BatchEngine *batch_engine_create(hipStream_t user_stream) {
  return new BatchEngine(user_stream);
}
void batch_engine_destroy(BatchEngine *e) { delete e; }
int64_t batch_engine_add(BatchEngine *e, const BatchedFit &f) {
  return e->add(f, true);
}
int64_t batch_engine_step(BatchEngine *e, int iterations, bool use_graph) {
  return e->step(iterations, use_graph);
}
std::vector<int64_t> batch_engine_take_finished(BatchEngine *e) {
  return e->take_finished();
}
int64_t batch_engine_resident_bytes(const BatchEngine *e) { return e->resident_bytes(); }
int64_t batch_engine_launches(const BatchEngine *e) { return e->launches_; }
void batch_engine_set_mixed(BatchEngine *e, bool mixed) {
  if (e->mixed_ == mixed) return;
  e->mixed_ = mixed;
  e->drop_graph();    // the captured iteration launches the other kernels
  e->dirty_ = true;
}

void batched_lloyd(const std::vector<BatchedFit> &in, int graph_replays,
                   hipStream_t user_stream, BatchedStats *stats, bool mixed) {
  if (in.empty()) return;
  // one allocation for all fits' workspaces (declared before the engine, so
  // it is released after the engine has drained its stream)
  std::vector<size_t> off(in.size());
  size_t total = 0;
  for (size_t i = 0; i < in.size(); ++i) {
    require(batched_fit_supported(in[i].n, in[i].d, in[i].k),
            "batched fit: unsupported shape");
    off[i] = total;
    total += (BatchEngine::arena_bytes(in[i]) + 255) & ~size_t(255);
  }
  TempBuffer pool(total, user_stream);
  BatchEngine eng(user_stream);
  int64_t max_iter = 0;
  for (size_t i = 0; i < in.size(); ++i) {
    eng.add(in[i], false, true, pool.as<char>() + off[i]);
    max_iter = std::max(max_iter, in[i].max_iter);
  }
  const int R = std::max(1, graph_replays);
  int64_t replays = 0;
  int64_t in_set = (int64_t)in.size();
  eng.defer_epilogue_ = true;
  eng.route_wide_ = true;
  eng.mixed_ = mixed;
  eng.tune_shape_ = true;
  // a fit with max_iter 0 (or converged) finishes at the first check
  for (;;) {
    const int64_t r = std::min<int64_t>(R, std::max<int64_t>(0, max_iter - replays));
    const int64_t active = eng.step((int)r, true);
    replays += r;
    if (active == 0) break;
    // once half the set has stopped, drop the stopped fits (their results are
    // written) so the work lists, grids and the captured iteration cover only
    // the fits still running
    if (2 * active <= in_set) {
      eng.flush_stopped();
      (void)eng.take_finished();
      in_set = active;
    }
  }
  eng.flush_stopped();
  // fits whose data range needs floating-point sums: the single-fit driver
  for (const BatchedFit &f : eng.wide_) driver_lloyd_single(f, user_stream);
  if (stats) {
    stats->graph_replays = replays;
    stats->assign_shape = eng.assign_shape_;
    stats->shape_ms[0] = eng.shape_ms_[0];
    stats->shape_ms[1] = eng.shape_ms_[1];
  }
  (void)hipStreamSynchronize(user_stream);
}

}  // namespace kmd
