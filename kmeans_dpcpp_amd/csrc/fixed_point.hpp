// Fixed-point scales of the exact centroid sums, shared by the single-fit and
// batched kernels so that both quantise every sample identically.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace kmd {

// maxbits[f] = max over samples of (bit pattern of X[f, s]) & 0x7fffffff
// (f < d), maxbits[d] the same for the weights.  scale[f] = 2^E_f with
// max|w| max|x_f| 2^E_f < 2^30 (each rounded element fits an int32) and
// n max|w| max|x_f| 2^E_f < 2^61 (the sum fits an int64); scale[d] likewise
// for the counts.  *nonfinite = 1 if any value is infinite or NaN.
__device__ inline void fixed_point_scales_from_bits(int64_t n, int d,
                                                    const unsigned *maxbits,
                                                    float *scale, int *nonfinite) {
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
  nonfinite[0] = bad;
}

}  // namespace kmd
