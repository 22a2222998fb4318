// Fixed-point scales of the exact centroid sums, shared by the single-fit and
// batched kernels so that both quantise every sample identically.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace kmd {

// Layout of the statistics gathered over X and w before a fit:
//   maxbits[f], f < d : max over samples of (bit pattern of X[f, s]) & 0x7fffffff
//   maxbits[d]        : the same for the weights
//   maxbits[d + 1]    : max over non-zero weights of ~(bits & 0x7fffffff), i.e.
//                       the complement of the smallest non-zero |w| (0: none)
//   estat[2 f], estat[2 f + 1] (f <= d): sum and count of the exponent fields
//                       of the row's normal (non-zero, finite) values
inline constexpr int64_t range_maxbits_count(int64_t d) { return d + 2; }
inline constexpr int64_t range_estat_count(int64_t d) { return 2 * (d + 1); }

// Adds one value's bit pattern to a thread's running maximum and exponent
// sum / count.
__device__ inline void range_accumulate(unsigned bits, unsigned &m, unsigned &esum,
                                        unsigned &ecnt) {
  const unsigned b = bits & 0x7fffffffu;
  m = b > m ? b : m;
  const unsigned e = b >> 23;
  const bool normal = e - 1u < 254u;
  esum += normal ? e : 0u;
  ecnt += normal ? 1u : 0u;
}

// Largest spread, in binary orders of magnitude, between the largest value
// of a row and the geometric mean of its normal values, beyond which the
// fixed-point sums are not used: with the step 2^-30 of the largest
// |w x_f|, a value of typical size keeps at least 30 - 12 = 18 bits.
constexpr double kMaxTypicalSpread = 12.0;
// Largest spread between the largest and the smallest non-zero weight: every
// weight (and so every count) keeps at least 30 - 13 = 17 bits.
constexpr int kMaxWeightSpread = 13;

// scale[f] = 2^E_f with max|w| max|x_f| 2^E_f < 2^30 (each rounded element
// fits an int32) and n max|w| max|x_f| 2^E_f < 2^61 (the sum fits an int64);
// scale[d] likewise for the counts.
// *flag = 1 if any value is infinite or NaN; else 2 if the data's dynamic
// range would leave typical elements (or the smallest weight) coarsely
// quantised (spreads above kMaxTypicalSpread / kMaxWeightSpread; estat may be
// NULL to skip this test); else 0.  Kernels sum in floating point whenever
// *flag != 0.
__device__ inline void fixed_point_scales_from_bits(
    int64_t n, int d, const unsigned *maxbits, const unsigned long long *estat,
    float *scale, int *flag) {
  const unsigned wb = maxbits[d];
  int bad = wb >= 0x7f800000u;
  const double wmax = (double)__uint_as_float(wb);
  for (int f = 0; f <= d; ++f) {
    const unsigned b = maxbits[f];
    bad |= b >= 0x7f800000u;
    const double m = wmax * (f < d ? (double)__uint_as_float(b) : 1.0);
    int e = 0;
    if (m > 0.0 && m < 1e300) {
      int ex_elem, ex_sum;
      frexp(m, &ex_elem);                       // m < 2^ex_elem
      frexp(m * (double)(n > 0 ? n : 1), &ex_sum);
      const int e1 = 30 - ex_elem, e2 = 61 - ex_sum;
      e = e1 < e2 ? e1 : e2;
      e = e > 120 ? 120 : (e < -120 ? -120 : e);
    }
    scale[f] = ldexpf(1.f, e);
  }
  int coarse = 0;
  if (estat != nullptr && !bad) {
    auto spread = [&](int f) {
      const unsigned long long cnt = estat[2 * f + 1];
      if (cnt == 0) return 0.0;
      const double mean_e = (double)estat[2 * f] / (double)cnt;
      return (double)(maxbits[f] >> 23) - mean_e;
    };
    double sx = 0.0;
    for (int f = 0; f < d; ++f) {
      const double sf = spread(f);
      sx = sf > sx ? sf : sx;
    }
    coarse |= sx + spread(d) > kMaxTypicalSpread;
    const unsigned wmin_inv = maxbits[d + 1];
    if (wmin_inv != 0u) {
      const int emin = (int)((~wmin_inv & 0x7fffffffu) >> 23);
      coarse |= (int)(wb >> 23) - emin > kMaxWeightSpread;
    }
  }
  flag[0] = bad ? 1 : (coarse ? 2 : 0);
}

}  // namespace kmd
