/*
 * C restatement of the reference's Lloyd hot path (TEST INFRASTRUCTURE ONLY:
 * the checker and the CPU baseline; see oracle/__init__.py).
 *
 * Arithmetic follows the reference kernels exactly where they define it:
 *  - dot[c, s] accumulates C[f, c] * X[f, s] with a fused multiply-add in
 *    ascending f (src/device_functions.hpp:120-136; the reference is built
 *    with icpx whose default floating-point model contracts `+= a * b` into
 *    an FMA);
 *  - pseudo-inertia = hn[c] - dot[c, s]; a centroid replaces the running
 *    best iff its value is strictly smaller, centroids scanned in ascending
 *    order from (0, +inf) (src/device_functions.hpp:139-162);
 *  - hn[c] = (sum_f C[f, c]^2) / 2 with FMA accumulation, f ascending
 *    (src/util_kernels.hpp:60-79);
 *  - the driver loop of src/kmeans_lloyd_driver.hpp:83-354 with the
 *    relocation order (sq-dist descending, index ascending) that this build
 *    defines for src/util_kernels.hpp:353.
 * Samples are processed in blocks of KMO_SB so the sample loop vectorises;
 * the per-(sample, centroid) operation sequence is unchanged by that.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define KMO_SB 32

#define DEFINE_KMO(T, SUF, FMA)                                                \
  void kmo_half_norm_##SUF(int64_t d, int64_t k, const T *C, T *hn) {          \
    for (int64_t c = 0; c < k; ++c) {                                          \
      T acc = (T)0;                                                            \
      for (int64_t f = 0; f < d; ++f) acc = FMA(C[f * k + c], C[f * k + c], acc); \
      hn[c] = acc / (T)2;                                                      \
    }                                                                          \
  }                                                                            \
                                                                               \
  static void assign_block_##SUF(int64_t n, int64_t d, int64_t k,              \
                                 const T *X, const T *C, const T *hn,          \
                                 int64_t s0, int64_t *lab, T *xb) {            \
    int64_t sb = n - s0 < KMO_SB ? n - s0 : KMO_SB;                            \
    for (int64_t f = 0; f < d; ++f)                                            \
      for (int64_t q = 0; q < KMO_SB; ++q)                                     \
        xb[f * KMO_SB + q] = q < sb ? X[f * n + s0 + q] : (T)0;                \
    T best[KMO_SB];                                                            \
    int64_t bidx[KMO_SB];                                                      \
    for (int64_t q = 0; q < KMO_SB; ++q) {                                     \
      best[q] = (T)INFINITY;                                                   \
      bidx[q] = 0;                                                             \
    }                                                                          \
    for (int64_t c = 0; c < k; ++c) {                                          \
      T acc[KMO_SB];                                                           \
      for (int64_t q = 0; q < KMO_SB; ++q) acc[q] = (T)0;                      \
      for (int64_t f = 0; f < d; ++f) {                                        \
        const T cv = C[f * k + c];                                             \
        const T *xr = xb + f * KMO_SB;                                         \
        _Pragma("omp simd") for (int64_t q = 0; q < KMO_SB; ++q)               \
            acc[q] = FMA(cv, xr[q], acc[q]);                                   \
      }                                                                        \
      const T h = hn[c];                                                       \
      for (int64_t q = 0; q < KMO_SB; ++q) {                                   \
        T v = h - acc[q];                                                      \
        int upd = v < best[q];                                                 \
        best[q] = upd ? v : best[q];                                           \
        bidx[q] = upd ? c : bidx[q];                                           \
      }                                                                        \
    }                                                                          \
    for (int64_t q = 0; q < sb; ++q) lab[s0 + q] = bidx[q];                    \
  }                                                                            \
                                                                               \
  void kmo_assign_##SUF(int64_t n, int64_t d, int64_t k, const T *X,           \
                        const T *C, const T *hn, int64_t *labels,              \
                        int nthreads) {                                        \
    int64_t nb = (n + KMO_SB - 1) / KMO_SB;                                    \
    _Pragma("omp parallel num_threads(nthreads)") {                            \
      T *xb = (T *)malloc(sizeof(T) * d * KMO_SB);                             \
      _Pragma("omp for schedule(static)") for (int64_t b = 0; b < nb; ++b)     \
          assign_block_##SUF(n, d, k, X, C, hn, b * KMO_SB, labels, xb);       \
      free(xb);                                                                \
    }                                                                          \
  }                                                                            \
                                                                               \
  /* Full fit; C holds init on entry and the final centroids on return. */    \
  int64_t kmo_lloyd_##SUF(int64_t n, int64_t d, int64_t k, const T *X,         \
                          const T *w, T *C, int64_t max_iter, double tol,      \
                          int64_t *labels, double *inertia, int nthreads) {    \
    T *hn = (T *)malloc(sizeof(T) * k);                                        \
    T *sums = (T *)malloc(sizeof(T) * d * k);                                  \
    T *sizes = (T *)malloc(sizeof(T) * k);                                     \
    T *newc = (T *)malloc(sizeof(T) * d * k);                                  \
    int64_t it = 0;                                                            \
    T shift = (T)INFINITY;                                                     \
    const T tol_t = (T)tol;                                                    \
    while (it < max_iter && shift > tol_t) {                                   \
      kmo_half_norm_##SUF(d, k, C, hn);                                        \
      kmo_assign_##SUF(n, d, k, X, C, hn, labels, nthreads);                   \
      memset(sums, 0, sizeof(T) * d * k);                                      \
      memset(sizes, 0, sizeof(T) * k);                                         \
      _Pragma("omp parallel num_threads(nthreads)") {                          \
        T *ps = (T *)calloc(d * k + k, sizeof(T));                             \
        _Pragma("omp for schedule(static)") for (int64_t s = 0; s < n; ++s) {  \
          int64_t l = labels[s];                                               \
          for (int64_t f = 0; f < d; ++f) ps[f * k + l] += X[f * n + s] * w[s]; \
          ps[d * k + l] += w[s];                                               \
        }                                                                      \
        _Pragma("omp critical") {                                              \
          for (int64_t i = 0; i < d * k; ++i) sums[i] += ps[i];                \
          for (int64_t c = 0; c < k; ++c) sizes[c] += ps[d * k + c];           \
        }                                                                      \
        free(ps);                                                              \
      }                                                                        \
      int64_t n_empty = 0;                                                     \
      for (int64_t c = 0; c < k; ++c) n_empty += sizes[c] == (T)0;             \
      if (n_empty > 0) {                                                       \
        /* squared distance to the current nearest centroid, then move the  \
           farthest samples (dist desc, index asc) into empty clusters in    \
           ascending cluster order */                                        \
        T *sq = (T *)malloc(sizeof(T) * n);                                    \
        int64_t *taken = (int64_t *)malloc(sizeof(int64_t) * n_empty);         \
        for (int64_t s = 0; s < n; ++s) {                                      \
          T acc = (T)0;                                                        \
          for (int64_t f = 0; f < d; ++f) {                                    \
            T df = X[f * n + s] - C[f * k + labels[s]];                        \
            acc = FMA(df, df, acc);                                            \
          }                                                                    \
          sq[s] = acc;                                                         \
        }                                                                      \
        int64_t m = n_empty < n ? n_empty : n;                                 \
        for (int64_t j = 0; j < m; ++j) {                                      \
          int64_t bs = -1;                                                     \
          for (int64_t s = 0; s < n; ++s) {                                    \
            int used = 0;                                                      \
            for (int64_t q = 0; q < j; ++q) used |= taken[q] == s;            \
            if (used) continue;                                                \
            if (bs < 0 || sq[s] > sq[bs]) bs = s;                              \
          }                                                                    \
          taken[j] = bs;                                                       \
        }                                                                      \
        int64_t j = 0;                                                         \
        for (int64_t e = 0; e < k && j < m; ++e) {                             \
          if (sizes[e] != (T)0) continue;                                      \
          int64_t s = taken[j++], p = labels[s];                               \
          for (int64_t f = 0; f < d; ++f) {                                    \
            T add = X[f * n + s] * w[s];                                       \
            sums[f * k + p] -= add;                                            \
            sums[f * k + e] = add;                                             \
          }                                                                    \
          sizes[p] -= w[s];                                                    \
          sizes[e] = w[s];                                                     \
        }                                                                      \
        free(sq);                                                              \
        free(taken);                                                           \
      }                                                                        \
      double sh = 0.0;                                                         \
      for (int64_t c = 0; c < k; ++c) {                                        \
        T acc = (T)0;                                                          \
        for (int64_t f = 0; f < d; ++f) {                                      \
          newc[f * k + c] = sums[f * k + c] / sizes[c];                        \
          T df = C[f * k + c] - newc[f * k + c];                               \
          acc = FMA(df, df, acc);                                              \
        }                                                                      \
        sh += (double)acc;                                                     \
      }                                                                        \
      memcpy(C, newc, sizeof(T) * d * k);                                      \
      shift = (T)sh;                                                           \
      ++it;                                                                    \
    }                                                                          \
    kmo_half_norm_##SUF(d, k, C, hn);                                          \
    kmo_assign_##SUF(n, d, k, X, C, hn, labels, nthreads);                     \
    double tot = 0.0;                                                          \
    for (int64_t s = 0; s < n; ++s) {                                          \
      T acc = (T)0;                                                            \
      for (int64_t f = 0; f < d; ++f) {                                        \
        T df = X[f * n + s] - C[f * k + labels[s]];                            \
        acc = FMA(df, df, acc);                                                \
      }                                                                        \
      tot += (double)(acc * w[s]);                                             \
    }                                                                          \
    *inertia = tot;                                                            \
    free(hn);                                                                  \
    free(sums);                                                                \
    free(sizes);                                                               \
    free(newc);                                                                \
    return it;                                                                 \
  }

DEFINE_KMO(float, f32, fmaf)
DEFINE_KMO(double, f64, fma)

int kmo_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
