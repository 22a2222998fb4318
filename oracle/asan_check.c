/* Sanitizer harness of the C restatement (oracle/restate.c).
 *
 * TEST INFRASTRUCTURE ONLY: built by `make -f oracle/sanitize.mk` twice from the
 * same sources, once with AddressSanitizer + UndefinedBehaviorSanitizer and
 * once plain, and run by tests/test_oracle_sanitized.py, which requires a
 * clean sanitizer run and identical output from the two builds.
 *
 * Every array is allocated at its exact size, so a read or write one
 * element past any of them is reported.  The cases cover the shapes where
 * index arithmetic is most likely to slip: n not a multiple of the 32-sample
 * block, d = 1, k = 1, n < k (clusters left empty after relocation),
 * centroids far from the data (empty clusters relocated on the first
 * iteration), max_iter = 0, float32 and float64.  Single-thread cases print
 * a digest of the labels and centroids; two-thread cases only have to run
 * clean (their per-thread partial sums are merged in arrival order). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void kmo_half_norm_f32(int64_t d, int64_t k, const float *C, float *hn);
void kmo_assign_f32(int64_t n, int64_t d, int64_t k, const float *X,
                    const float *C, const float *hn, int64_t *labels,
                    int nthreads);
int64_t kmo_lloyd_f32(int64_t n, int64_t d, int64_t k, const float *X,
                      const float *w, float *C, int64_t max_iter, double tol,
                      int64_t *labels, double *inertia, int nthreads);
void kmo_half_norm_f64(int64_t d, int64_t k, const double *C, double *hn);
void kmo_assign_f64(int64_t n, int64_t d, int64_t k, const double *X,
                    const double *C, const double *hn, int64_t *labels,
                    int nthreads);
int64_t kmo_lloyd_f64(int64_t n, int64_t d, int64_t k, const double *X,
                      const double *w, double *C, int64_t max_iter, double tol,
                      int64_t *labels, double *inertia, int nthreads);

static uint64_t g_state = 0x9e3779b97f4a7c15ull;

/* uniform in [0, 1) from a 64-bit LCG (Knuth's MMIX constants) */
static double urand(void) {
  g_state = g_state * 6364136223846793005ull + 1442695040888963407ull;
  return (double)(g_state >> 11) * (1.0 / 9007199254740992.0);
}

/* FNV-1a over bytes */
static uint64_t digest(const void *p, size_t bytes, uint64_t h) {
  const unsigned char *b = (const unsigned char *)p;
  for (size_t i = 0; i < bytes; ++i) h = (h ^ b[i]) * 0x100000001b3ull;
  return h;
}

#define DEFINE_CASE(T, SUF)                                                    \
  static void case_##SUF(int64_t n, int64_t d, int64_t k, int64_t max_iter,    \
                         int far, int nthreads) {                              \
    T *X = (T *)malloc(sizeof(T) * (size_t)(d * n));                           \
    T *w = (T *)malloc(sizeof(T) * (size_t)n);                                 \
    T *C = (T *)malloc(sizeof(T) * (size_t)(d * k));                           \
    T *hn = (T *)malloc(sizeof(T) * (size_t)k);                                \
    int64_t *lab = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);             \
    /* samples around 4 centres, weights in [0.5, 1.5) */                      \
    for (int64_t s = 0; s < n; ++s) {                                          \
      const int64_t g = s % 4;                                                 \
      for (int64_t f = 0; f < d; ++f)                                          \
        X[f * n + s] = (T)(3.0 * (double)((g + f) % 4) + urand() - 0.5);       \
      w[s] = (T)(0.5 + urand());                                               \
    }                                                                          \
    /* init: samples c mod n, the first `far` of them moved far away */        \
    for (int64_t c = 0; c < k; ++c)                                            \
      for (int64_t f = 0; f < d; ++f)                                          \
        C[f * k + c] = X[f * n + c % n] + (c < far ? (T)1000 : (T)0);          \
    kmo_half_norm_##SUF(d, k, C, hn);                                          \
    kmo_assign_##SUF(n, d, k, X, C, hn, lab, nthreads);                        \
    uint64_t h0 = digest(lab, sizeof(int64_t) * (size_t)n, 0xcbf29ce484222325ull); \
    double inertia = 0.0;                                                      \
    const int64_t it = kmo_lloyd_##SUF(n, d, k, X, w, C, max_iter, 0.0, lab,   \
                                       &inertia, nthreads);                    \
    if (nthreads == 1) {                                                       \
      uint64_t h1 = digest(lab, sizeof(int64_t) * (size_t)n, h0);              \
      h1 = digest(C, sizeof(T) * (size_t)(d * k), h1);                         \
      printf("%s n=%lld d=%lld k=%lld max_iter=%lld far=%d: iters %lld "       \
             "inertia %.17g digest %016llx\n",                                 \
             #SUF, (long long)n, (long long)d, (long long)k,                   \
             (long long)max_iter, far, (long long)it, inertia,                 \
             (unsigned long long)h1);                                          \
    }                                                                          \
    free(X);                                                                   \
    free(w);                                                                   \
    free(C);                                                                   \
    free(hn);                                                                  \
    free(lab);                                                                 \
  }

DEFINE_CASE(float, f32)
DEFINE_CASE(double, f64)

int main(void) {
  static const int64_t shapes[][5] = {
      /* n, d, k, max_iter, far */
      {1000, 5, 7, 10, 0},  {1000, 5, 7, 10, 3}, {33, 1, 1, 4, 0},
      {31, 2, 3, 0, 0},     {5, 3, 8, 3, 0},     {200, 4, 6, 6, 2},
      {64, 33, 40, 3, 5},   {1, 2, 3, 2, 1},
  };
  const int ns = (int)(sizeof(shapes) / sizeof(shapes[0]));
  for (int threads = 1; threads <= 2; ++threads)
    for (int i = 0; i < ns; ++i) {
      const int64_t *s = shapes[i];
      case_f32(s[0], s[1], s[2], s[3], (int)s[4], threads);
      case_f64(s[0], s[1], s[2], s[3], (int)s[4], threads);
    }
  return 0;
}
