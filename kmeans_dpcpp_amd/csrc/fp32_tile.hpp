// The FP32 matrix-core tile arithmetic shared by every kernel that must give
// the FP32 path's labels bit for bit (the fused step, its re-score modes and
// the in-kernel re-score of the mixed-precision step): one 32-centroid tile
// of v_mfma_f32_32x32x2f32 products per 32-sample column block, the chain
// started from the negated half norms, and the first-minimum scan in (tile,
// row) order.
#pragma once
#include <hip/hip_runtime.h>

namespace kmd {
namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned long long u64x8 __attribute__((ext_vector_type(8)));

// acc[q] = -hn[tile] + C[:, tile] . X[:, columns of block q] for one
// 32-centroid tile: the chain starts from the negated half norms (the C
// operand of the first MFMA), so -acc is the score hn - c.x with no vector
// instruction per score.  Ct = this lane's A operands of the tile:
// Ct[64 ks] = C[2 ks + half][col] (the kernels' LDS layout), so every k-step
// reads at an immediate offset; Ht = the tile's -hn at this lane's first row.
template <int KS, int S>
__device__ inline void tile_mfma(const float *__restrict__ Ct,
                                 const float *__restrict__ Ht,
                                 const float (&xr)[KS][S], f32x16 (&acc)[S]) {
  f32x16 h;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const float4 h4 = *reinterpret_cast<const float4 *>(Ht + 8 * g);
    h[4 * g + 0] = h4.x;
    h[4 * g + 1] = h4.y;
    h[4 * g + 2] = h4.z;
    h[4 * g + 3] = h4.w;
  }
  const float a0 = Ct[0];
#pragma unroll
  for (int q = 0; q < S; ++q)
    acc[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, xr[0][q], h, 0, 0, 0);
#pragma unroll
  for (int ks = 1; ks < KS; ++ks) {
    const float a = Ct[64 * ks];
#pragma unroll
    for (int q = 0; q < S; ++q)
      acc[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, xr[ks][q], acc[q], 0, 0,
                                                     0);
  }
}

// Argmin over the 16 rows this lane holds in tile `ct` (row r is centroid
// 32 ct + 8 (r / 4) + 4 half + r % 4, ascending in r), merged into the
// running best: the first minimum in (tile, row) order, so ties keep the
// lowest centroid, as one running scan in that order would.  Rows are
// tracked as r (an inline constant) and tiles as ct; the centroid index is
// formed once at the end (best_index).
template <int S>
__device__ inline void tile_argmin(int ct, const f32x16 (&acc)[S],
                                   float (&bestv)[S], int (&bestt)[S],
                                   int (&bestr)[S]) {
  float tv[S];
  int tr[S];
#pragma unroll
  for (int q = 0; q < S; ++q) {
    tv[q] = __builtin_huge_valf();
    tr[q] = 0;
  }
#pragma unroll
  for (int g = 0; g < 4; ++g) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
#pragma unroll
      for (int q = 0; q < S; ++q) {
        const float v = -acc[q][4 * g + e];   // hn - c.x (exact negation)
        const bool upd = v < tv[q];
        tv[q] = upd ? v : tv[q];
        tr[q] = upd ? (4 * g + e) : tr[q];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < S; ++q) {
    const bool upd = tv[q] < bestv[q];
    bestv[q] = upd ? tv[q] : bestv[q];
    bestt[q] = upd ? ct : bestt[q];
    bestr[q] = upd ? tr[q] : bestr[q];
  }
}

// v_max3_f32, written out so that no canonicalisation is put in front of
// it: NaN operands are passed over (IEEE mode), so a NaN score never wins.
__device__ inline float vmax3(float a, float b, float c) {
  float r;
  asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// Largest of a lane's 16 accumulators (NaNs passed over): a max3 tree of
// depth 3 rather than a chain of 8.  The tree's last operand is -inf, so a
// lane whose 16 rows are all NaN reports -inf (no candidate) rather than
// NaN: the scan of the reference skips NaN scores, and a NaN here would
// otherwise win the merge with the other half of the column.
__device__ inline float tile_max(const f32x16 &acc) {
  const float t0 = vmax3(acc[0], acc[1], acc[2]);
  const float t1 = vmax3(acc[3], acc[4], acc[5]);
  const float t2 = vmax3(acc[6], acc[7], acc[8]);
  const float t3 = vmax3(acc[9], acc[10], acc[11]);
  const float t4 = vmax3(acc[12], acc[13], acc[14]);
  const float u0 = vmax3(t0, t1, t2);
  const float u1 = vmax3(t3, t4, acc[15]);
  return vmax3(u0, u1, -__builtin_huge_valf());
}

// The same first minimum in (tile, row) order as tile_argmin, with less
// vector work per tile: the tile's best (largest acc = smallest score) by a
// max3 chain, and on a strict improvement the tile index and a copy of its
// 16 accumulators; the row is resolved once per sample by best_row.  The
// copy is eight v_mov_b64 under the lanes' improvement mask rather than 16
// selects (the compiler splits a 64-bit copy back into 32-bit moves, hence
// the asm); it runs after tile_max has read the same registers, so the
// MFMA chain's results are complete by then.
template <int S>
__device__ inline void tile_track_best(int ct, const f32x16 (&acc)[S],
                                       float (&bestm)[S], int (&bestt)[S],
                                       f32x16 (&besta)[S]) {
#pragma unroll
  for (int q = 0; q < S; ++q) {
    const float m = tile_max(acc[q]);
    const bool upd = m > bestm[q];   // ties keep the earlier tile
    bestm[q] = upd ? m : bestm[q];
    bestt[q] = upd ? ct : bestt[q];
    u64x8 b = __builtin_bit_cast(u64x8, besta[q]);
    const u64x8 a = __builtin_bit_cast(u64x8, acc[q]);
    if (upd) {
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        unsigned long long v;
        asm volatile("v_mov_b64 %0, %1" : "=v"(v) : "v"(a[i]));
        b[i] = v;
      }
    }
    besta[q] = __builtin_bit_cast(f32x16, b);
  }
}

// Lowest row of the kept tile that holds the best value.
__device__ inline int best_row(const f32x16 &besta, float bestm) {
  int r0 = 0;
#pragma unroll
  for (int r = 15; r >= 0; --r) r0 = (besta[r] == bestm) ? r : r0;
  return r0;
}

__device__ inline int best_index(int t, int r, int half) {
  return t * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
}

}  // namespace
}  // namespace kmd
