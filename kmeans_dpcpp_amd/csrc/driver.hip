// Host driver of a full Lloyd fit (src/kmeans_lloyd_driver.hpp:21-431).
//
// Semantics kept from the reference: iterate while it < max_iter and the sum
// of squared centroid shifts, compared in the data type, exceeds tol (raw, not
// variance-scaled); each iteration computes half-norms, the fused
// assignment+update, the reduction of private copies with the empty-cluster
// list, the relocation of empty clusters (on unnormalised sums, before the
// division), the division, the shifts and their sum; the epilogue recomputes
// labels and the weighted inertia of the final centroids and copies them to
// res_centroids_t.  init_centroids_t is one of the rotating centroid buffers
// and is overwritten.
//
// Differences (all on the host side, none in results):
//  * the host is off the GPU's critical path: the division, shifts and
//    shift sum run in one kernel that skips itself when the iteration found
//    empty clusters, so a whole iteration is enqueued without a host
//    decision, and iteration i + 1 is enqueued before iteration i's scalars
//    {shift sum, empty count} are read back (one wait per iteration, on an
//    event, while the GPU runs the next iteration).  The work enqueued ahead
//    is discarded when the fit stops, or redone after a relocation; the
//    centroid buffers rotate over three (init, res, a scratch one) so it
//    never overwrites centroids still needed;
//  * private copies are one per block of the persistent fused kernel and are
//    written with plain stores, so the sums are deterministic;
//  * the relocation path reuses the labels of the fused step (kept per
//    iteration parity) instead of recomputing them from the same centroids
//    and half-norms (src/kmeans_lloyd_driver.hpp:226-250);
//  * the verbose line reports the inertia of the centroids that produced the
//    labels (the reference evaluates it against the undivided sums).
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "lloyd_fast.hpp"

namespace kmd {

template <typename T, typename I>
struct LloydWorkspace {
  int64_t n, d, k;
  int blocks = 1;          // private copies / partial slots
  bool fast = false;       // MFMA fast path (fixed-point partials)
  bool store_flush = false;
  int mixed_terms = 0;     // BF16 split terms (1 or 3) or 8 (FP8) of the mixed path, 0: FP32
  bool tiled = false;      // tiled path: centroids streamed, separate update
  int blocks_b = 0;        // partial slots of the FP32 re-score (mixed path)
  // LDS-resident mixed path: the re-score restricted to candidate tiles,
  // from the sample-major copy xr, and the queue's separate exact update
  bool mixed_cand = false;
  // LDS-resident mixed path: undecided samples re-scored inside the kernel
  bool mixed_ink = false;
  FusedPlan plan;
  TempBuffer hn, sizes, shifts, per_sample, copies, sizes_copies, list,
      scalars, scales;
  TempBuffer aop{0, nullptr}, bound{0, nullptr}, queue{0, nullptr},
      queue_counts{0, nullptr};
  TempBuffer xr{0, nullptr}, upd{0, nullptr};  // tiled: X sample-major, update scratch
  TempBuffer qmask{0, nullptr}, cand{0, nullptr};  // tiled BF16: candidate tiles
  TempBuffer xb{0, nullptr}, xnb{0, nullptr};  // tiled BF16 (T = 1): BF16 samples
  int64_t seg_cap = 0;     // capacity of each block's re-score queue segment
  // Two iterations can be in flight (the next one is enqueued before this
  // one's scalars are read back): cluster sizes, the empty list and the
  // scalars {shift sum @0, n_empty @8, re-scored count @16} have one slot per
  // iteration parity `par`.  cbuf: the third centroid buffer of the rotation.
  int par = 0;
  TempBuffer cbuf, fin_part, fin_ticket;
  // copies_bytes / sizes_bytes: private copies of the generic kernel, or the
  // fast path's per-block partials (in `copies`, sizes_bytes = 0)
  LloydWorkspace(int64_t n_, int64_t d_, int64_t k_, int blocks_, bool fast_,
                 bool store_, const FusedPlan &plan_, size_t copies_bytes,
                 size_t sizes_bytes, hipStream_t s)
      : n(n_), d(d_), k(k_), blocks(blocks_), fast(fast_), store_flush(store_),
        plan(plan_), hn(sizeof(T) * k_, s), sizes(2 * sizeof(T) * k_, s),
        shifts(sizeof(T) * k_, s), per_sample(sizeof(T) * (n_ > 0 ? n_ : 1), s),
        copies(copies_bytes, s), sizes_copies(sizes_bytes, s),
        list(2 * sizeof(I) * (k_ + 1), s), scalars(64, s),
        scales(fast_ ? sizeof(float) * (d_ + 1) + 16 : 0, s),
        cbuf(sizeof(T) * d_ * k_, s),
        fin_part(sizeof(double) * finalize_step_blocks(k_), s),
        fin_ticket(sizeof(unsigned), s) {
    KMD_HIP_CHECK(hipMemsetAsync(fin_ticket.as<unsigned>(), 0, sizeof(unsigned), s));
  }
  T *sizes_p() const { return sizes.as<T>() + par * k; }
  I *list_p() const { return list.as<I>() + par * (k + 1); }
  char *scalars_p() const { return scalars.as<char>() + 32 * par; }
  I *n_empty() const { return reinterpret_cast<I *>(scalars_p() + 8); }
  int *queue_count() const { return reinterpret_cast<int *>(scalars_p() + 16); }
  double *shift_sum() const { return reinterpret_cast<double *>(scalars_p()); }
  float *scale() const { return scales.as<float>(); }
  int *nonfinite() const {
    return reinterpret_cast<int *>(scales.as<char>() + sizeof(float) * (d + 1));
  }
};

// Re-score forms of the LDS-resident mixed path.  By default undecided
// samples are re-scored inside the BF16 kernel (when the FP32 centroid tiles
// fit LDS beside its own data), else by a second kernel that gathers every
// queued sample over all tiles and adds it to the sums.
// KMD_MIXED_GATHER=1 keeps the second kernel; KMD_MIXED_CAND=1 switches to
// the candidate-tile re-score from a sample-major copy of X plus a separate
// exact update of the queue (fewer bytes per queued sample, four more
// launches per iteration; at n = 10M, d = 32, k = 256: 630 iters/s against
// the gather's 795).
static bool env_is_1(const char *name) {
  const char *e = std::getenv(name);
  return e != nullptr && e[0] == '1';
}
static bool mixed_cand_requested() { return env_is_1("KMD_MIXED_CAND"); }
static bool mixed_gather_requested() { return env_is_1("KMD_MIXED_GATHER"); }

// Mixed-precision workspace: BF16 kernel partials, then FP32 re-score
// partials; returns nullptr when the shape is not served or the inputs hold
// non-finite values (the FP32 path handles those).
template <typename T, typename I>
static std::unique_ptr<LloydWorkspace<T, I>> make_mixed_workspace(
    int64_t n, int64_t d, int64_t k, int terms, const T *X, const T *w,
    hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    if (n <= 0 || !fast_fused_supported<T>(d, k)) return nullptr;
    // candidate tiles: the re-score of an undecided sample covers only the
    // 32-centroid tiles whose BF16 best is within the bound of its best,
    // reading the sample from a sample-major copy of X (made once per fit)
    const bool cand = tiled_cand_supported(k) && tiled_supported(d, k) &&
                      mixed_cand_fits(d, k, terms) && mixed_cand_requested();
    const bool ink = !cand && !mixed_gather_requested() && mixed_ink_fits(d, k, terms);
    const int ba = mixed_blocks<I>(n, d, k, terms, cand, ink);
    const int bb = ink ? 0 : fast_rescore_blocks<I>(d, k, ba);
    if (ba < 1 || (!ink && bb < 1)) return nullptr;
    auto ws = std::make_unique<LloydWorkspace<T, I>>(
        n, d, k, ba, true, true, FusedPlan{},
        sizeof(int64_t) * (size_t)(ba + bb) * (d * k + k), 0, s);
    ws->mixed_terms = terms;
    ws->mixed_ink = ink;
    ws->blocks_b = bb;
    ws->aop = TempBuffer(mixed_aop_bytes(d, k, terms), s);
    ws->bound = TempBuffer(2 * sizeof(float), s);
    ws->seg_cap = ink ? 0 : mixed_seg_cap(n, d, ba);
    ws->queue = TempBuffer(sizeof(int) * (size_t)ba * ws->seg_cap, s);
    ws->queue_counts = TempBuffer(sizeof(int) * (size_t)ba, s);
    launch_fixed_point_scales(n, d, X, w, ws->scale(), ws->nonfinite(), s);
    int bad = 0;
    KMD_HIP_CHECK(hipMemcpyAsync(&bad, ws->nonfinite(), sizeof(int),
                                 hipMemcpyDeviceToHost, s));
    KMD_HIP_CHECK(hipStreamSynchronize(s));
    if (bad) return nullptr;
    if (cand) {
      ws->mixed_cand = true;
      ws->qmask = TempBuffer(sizeof(unsigned) * (size_t)ba * ws->seg_cap, s);
      ws->cand = TempBuffer(tiled_cand_scratch_bytes(n, k), s);
      ws->xr = TempBuffer(sizeof(float) * (size_t)n * d, s);
      launch_transpose(n, d, X, ws->xr.template as<float>(), s);
    }
    return ws;
  } else {
    (void)n, (void)d, (void)k, (void)terms, (void)X, (void)w, (void)s;
    return nullptr;
  }
}

// Tiled workspace (float, finite inputs): X_r, one int64 sum slot, update
// scratch; with precision != FP32 also the BF16 operands and queue.  Returns
// nullptr for non-finite inputs (the generic path propagates them).
template <typename T, typename I>
static std::unique_ptr<LloydWorkspace<T, I>> make_tiled_workspace(
    int64_t n, int64_t d, int64_t k, int precision, const T *X, const T *w,
    hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    if (n <= 0 || !tiled_supported(d, k)) return nullptr;
    auto ws = std::make_unique<LloydWorkspace<T, I>>(
        n, d, k, 1, true, true, FusedPlan{}, sizeof(int64_t) * (size_t)(d * k + k),
        0, s);
    launch_fixed_point_scales(n, d, X, w, ws->scale(), ws->nonfinite(), s);
    int bad = 0;
    KMD_HIP_CHECK(hipMemcpyAsync(&bad, ws->nonfinite(), sizeof(int),
                                 hipMemcpyDeviceToHost, s));
    KMD_HIP_CHECK(hipStreamSynchronize(s));
    if (bad) return nullptr;
    ws->tiled = true;
    ws->xr = TempBuffer(sizeof(float) * (size_t)n * d, s);
    launch_transpose(n, d, X, ws->xr.template as<float>(), s);
    ws->upd = TempBuffer(tiled_update_scratch_bytes(n, k), s);
    if (precision != KMD_PREC_FP32) {
      // BF16 terms, or 8: the FP8 (e4m3) kernel
      const int terms = precision == KMD_PREC_BF16X3 ? 3
                        : precision == KMD_PREC_FP8  ? 8
                                                     : 1;
      const int ba = tiled_mixed_blocks<I>(n, d, k, terms);
      if (ba >= 1) {
        ws->mixed_terms = terms;
        ws->blocks_b = ba;
        ws->seg_cap = tiled_mixed_seg_cap(n, d, terms, ba);
        ws->aop = TempBuffer(tiled_aop_bytes(d, k, terms), s);
        ws->bound = TempBuffer(sizeof(float) * tiled_bound_floats(), s);
        ws->queue = TempBuffer(sizeof(int) * (size_t)ba * ws->seg_cap, s);
        ws->queue_counts = TempBuffer(sizeof(int) * (size_t)ba, s);
        if (tiled_cand_supported(k)) {
          ws->qmask = TempBuffer(sizeof(unsigned) * (size_t)ba * ws->seg_cap, s);
          ws->cand = TempBuffer(tiled_cand_scratch_bytes(n, k), s);
        }
        if (terms == 1 || terms == 8) {
          ws->xb = TempBuffer(tiled_xb_bytes(n, d), s);
          ws->xnb = TempBuffer(sizeof(float) * (size_t)n, s);
          if (terms == 1)
            launch_tiled_xb_prep(n, d, X, ws->xb.template as<void>(),
                                 ws->xnb.template as<float>(), s);
          else
            launch_tiled_x8_prep(n, d, X, ws->xb.template as<void>(),
                                 ws->xnb.template as<float>(),
                                 ws->bound.template as<float>(), s);
        }
      }
    }
    return ws;
  } else {
    (void)n, (void)d, (void)k, (void)precision, (void)X, (void)w, (void)s;
    return nullptr;
  }
}

// ---- launch tuning ----------------------------------------------------------
// The choice in force on this thread (read by the planners while a driver
// call runs), the autotuner's table per (d, k, precision, log2 n) and a
// process-wide override (kmd_set_tuning).
namespace {
thread_local int t_blocks_cap = 0;
thread_local int t_path = 0;
thread_local int t_shape = 0;
struct TuneChoice {
  int path = 0, blocks_per_cu = 0, shape = 0;
};
struct TuneGuard {
  int pb, pp, ps;
  explicit TuneGuard(const TuneChoice &c)
      : pb(t_blocks_cap), pp(t_path), ps(t_shape) {
    t_blocks_cap = c.blocks_per_cu;
    t_path = c.path;
    t_shape = c.shape;
  }
  ~TuneGuard() {
    t_blocks_cap = pb;
    t_path = pp;
    t_shape = ps;
  }
};
std::mutex g_tune_mu;
std::map<std::tuple<int64_t, int64_t, int, int>, TuneChoice> g_tune;
bool g_override = false;
TuneChoice g_override_choice;

int log2_bucket(int64_t n) {
  int b = 0;
  while ((int64_t(1) << (b + 1)) <= n) ++b;
  return b;
}

TuneChoice choose_tuning(int64_t n, int64_t d, int64_t k, int precision) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (g_override) return g_override_choice;
  auto it = g_tune.find(std::make_tuple(d, k, precision, log2_bucket(n)));
  return it == g_tune.end() ? TuneChoice{} : it->second;
}
}  // namespace

int tuned_blocks_per_cu() { return t_blocks_cap; }
int tuned_path() { return t_path; }
int tuned_shape() { return t_shape; }

void set_tuning_override(int path, int blocks_per_cu, int shape) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_override = path >= 0;
  g_override_choice.path = path < 0 ? 0 : path;
  g_override_choice.blocks_per_cu = blocks_per_cu < 0 ? 0 : blocks_per_cu;
  g_override_choice.shape = shape < 0 ? 0 : shape;
}

// Test hook: KMD_FORCE_TILED=1 selects the tiled path for every shape it
// serves (the LDS-resident kernels are otherwise preferred).
static bool force_tiled() {
  const char *e = std::getenv("KMD_FORCE_TILED");
  return e != nullptr && e[0] == '1';
}

template <typename T, typename I>
static std::unique_ptr<LloydWorkspace<T, I>> make_workspace(int64_t n,
                                                            int64_t d,
                                                            int64_t k,
                                                            const T *X,
                                                            const T *w,
                                                            int precision,
                                                            hipStream_t s) {
  // With k <= 32 the step is bound by reading X, not by the cross term:
  // every precision runs the FP32 few-centroid kernel there (its labels and
  // sums are what the mixed paths reproduce, at a fraction of their time).
  const bool smallk = std::is_same<T, float>::value && !force_tiled() &&
                      tuned_path() != 2 && fast_smallk_serves(d, k);
  // the FP8 cross term is built on the tiled kernels only
  const bool tiled_first = !smallk && (force_tiled() || tuned_path() == 2 ||
                                       (precision == KMD_PREC_FP8 &&
                                        std::is_same<T, float>::value &&
                                        tiled_supported(d, k)));
  if (precision != KMD_PREC_FP32 && precision != KMD_PREC_FP8 && !tiled_first &&
      !smallk) {
    auto ws = make_mixed_workspace<T, I>(
        n, d, k, precision == KMD_PREC_BF16X3 ? 3 : 1, X, w, s);
    if (ws) return ws;
  }
  if (tiled_first || !fast_fused_supported<T>(d, k)) {
    auto ws = make_tiled_workspace<T, I>(n, d, k, precision, X, w, s);
    if (ws) return ws;
  }
  if (n > 0 && fast_fused_supported<T>(d, k)) {
    const int blocks = fast_fused_blocks<I>(n, d, k);
    auto ws = std::make_unique<LloydWorkspace<T, I>>(
        n, d, k, blocks, true, true, FusedPlan{},
        sizeof(int64_t) * (size_t)blocks * (d * k + k), 0, s);
    // fixed-point scales: X and w do not change during the fit
    if constexpr (std::is_same<T, float>::value)
      launch_fixed_point_scales(n, d, X, w, ws->scale(), ws->nonfinite(), s);
    return ws;
  }
  FusedPlan p = plan_generic_fused<T>(n, d, k);
  if (p.priv)
    return std::make_unique<LloydWorkspace<T, I>>(
        n, d, k, p.blocks, false, true, p, sizeof(T) * (size_t)p.blocks * d * k,
        sizeof(T) * (size_t)p.blocks * k, s);
  // sums too large for LDS: global atomics into a few zeroed copies
  const int copies = 8;
  return std::make_unique<LloydWorkspace<T, I>>(
      n, d, k, copies, false, false, p, sizeof(T) * (size_t)copies * d * k,
      sizeof(T) * (size_t)copies * k, s);
}

// Optional timing of the fused-step kernel with HIP events on the driver's
// stream: one event pair per iteration parity (two iterations can be in
// flight), read back once that iteration is known to have finished.
struct FusedTimer {
  bool on = false;
  hipEvent_t start[2] = {nullptr, nullptr}, stop[2] = {nullptr, nullptr};
  bool pending[2] = {false, false};
  int slot = 0;
  double total_ms = 0.0;
  int64_t launches = 0;
  explicit FusedTimer(bool enable) : on(enable) {
    if (on) {
      for (int i = 0; i < 2; ++i) {
        KMD_HIP_CHECK(hipEventCreate(&start[i]));
        KMD_HIP_CHECK(hipEventCreate(&stop[i]));
      }
    }
  }
  ~FusedTimer() {
    for (int i = 0; i < 2; ++i) {
      if (start[i]) (void)hipEventDestroy(start[i]);
      if (stop[i]) (void)hipEventDestroy(stop[i]);
    }
  }
  void begin(hipStream_t s) {
    if (on) KMD_HIP_CHECK(hipEventRecord(start[slot], s));
  }
  void end(hipStream_t s) {
    if (!on) return;
    KMD_HIP_CHECK(hipEventRecord(stop[slot], s));
    pending[slot] = true;
  }
  // call once the work timed in slot i has finished
  void collect(int i) {
    if (!on || !pending[i]) return;
    float ms = 0.f;
    KMD_HIP_CHECK(hipEventElapsedTime(&ms, start[i], stop[i]));
    total_ms += ms;
    ++launches;
    pending[i] = false;
  }
  void collect_all() {
    collect(0);
    collect(1);
  }
  void discard(int i) { pending[i] = false; }
};

// Pinned host slots for the per-iteration scalars (32 bytes per parity), so
// their device-to-host copies are asynchronous.  Pinning and unpinning cost
// hundreds of microseconds, so the 64-byte blocks are kept in a process-wide
// free list and reused (one per driver call running at a time).
std::mutex g_pinned_mu;
std::vector<char *> g_pinned_free;

struct PinnedScalars {
  char *p = nullptr;
  PinnedScalars() {
    {
      std::lock_guard<std::mutex> lk(g_pinned_mu);
      if (!g_pinned_free.empty()) {
        p = g_pinned_free.back();
        g_pinned_free.pop_back();
      }
    }
    if (p == nullptr)
      KMD_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&p), 64, hipHostMallocDefault));
    std::memset(p, 0, 64);
  }
  ~PinnedScalars() {
    if (p == nullptr) return;
    std::lock_guard<std::mutex> lk(g_pinned_mu);
    g_pinned_free.push_back(p);
  }
  PinnedScalars(const PinnedScalars &) = delete;
  PinnedScalars &operator=(const PinnedScalars &) = delete;
  char *slot(int i) const { return p + 32 * i; }
};

struct IterEvents {
  hipEvent_t e[2] = {nullptr, nullptr};
  IterEvents() {
    for (int i = 0; i < 2; ++i)
      KMD_HIP_CHECK(hipEventCreateWithFlags(&e[i], hipEventDisableTiming));
  }
  ~IterEvents() {
    for (int i = 0; i < 2; ++i)
      if (e[i]) (void)hipEventDestroy(e[i]);
  }
  IterEvents(const IterEvents &) = delete;
  IterEvents &operator=(const IterEvents &) = delete;
};

// LDS-resident mixed path: BF16 step (operands already prepared for `cur`)
// and the re-score of its undecided samples, in the workspace's form:
// labels (the FP32 path's) and the int64 partial slots.
template <typename T, typename I>
static void mixed_labels_and_partials(LloydWorkspace<T, I> &ws, const T *X,
                                      const T *w, const T *cur, I *labels,
                                      hipStream_t s, const int *skip) {
  if constexpr (std::is_same<T, float>::value) {
    const int64_t n = ws.n, d = ws.d, k = ws.k;
    launch_mixed_step<I>(n, d, k, ws.mixed_terms, ws.blocks, X, w,
                         ws.aop.template as<const void>(),
                         ws.bound.template as<const float>(), ws.scale(),
                         labels, ws.copies.template as<void>(),
                         ws.queue.template as<int>(),
                         ws.queue_counts.template as<int>(),
                         ws.mixed_cand ? ws.qmask.template as<unsigned>() : nullptr, s,
                         skip, ws.mixed_ink ? cur : nullptr,
                         ws.mixed_ink ? ws.hn.template as<const T>() : nullptr,
                         ws.mixed_ink ? ws.queue_count() : nullptr);
    if (ws.mixed_ink) {
      // re-scored inside the kernel: nothing more to launch
    } else if (ws.mixed_cand) {
      // FP32 re-score of the candidate tiles (labels), then the exact sums
      // of the re-scored samples into the re-score partial slots
      launch_tiled_cand_rescore<I>(
          n, d, k, ws.xr.template as<const float>(), cur, ws.hn.template as<const T>(),
          labels, ws.queue.template as<const int>(),
          ws.queue_counts.template as<const int>(),
          ws.qmask.template as<const unsigned>(), ws.blocks, ws.seg_cap,
          ws.cand.template as<void>(), ws.queue_count(), s);
      launch_queue_update<I>(d, k, ws.blocks_b, ws.blocks, ws.seg_cap,
                             ws.queue.template as<const int>(),
                             ws.queue_counts.template as<const int>(),
                             ws.xr.template as<const float>(), w, labels, ws.scale(),
                             ws.copies.template as<int64_t>() + (size_t)ws.blocks * (d * k + k),
                             s, skip);
    } else {
      launch_fast_rescore<I>(n, d, k, X, w, cur, ws.hn.template as<const T>(),
                             labels, ws.scale(), ws.nonfinite(),
                             ws.copies.template as<void>(), ws.blocks,
                             ws.queue.template as<const int>(),
                             ws.queue_counts.template as<const int>(),
                             ws.blocks, ws.seg_cap, ws.queue_count(), s);
    }
  } else {
    (void)ws, (void)X, (void)w, (void)cur, (void)labels, (void)s, (void)skip;
    throw RuntimeError("mixed-precision step: float32 only");
  }
}

// hn (unless the previous iteration's finalize computed it), fused step and
// reduction for centroids `cur`, sums into `next`; the empty count is
// written, not accumulated.
template <typename T, typename I>
static void lloyd_front(LloydWorkspace<T, I> &ws, const T *X, const T *w,
                        const T *cur, T *next, I *labels, hipStream_t s,
                        FusedTimer &timer, bool hn_ready, const int *skip) {
  const int64_t n = ws.n, d = ws.d, k = ws.k;
  if (!hn_ready) launch_half_l2_norm<T>(d, k, cur, ws.hn.template as<T>(), s);
  if constexpr (std::is_same<T, float>::value) {
    if (ws.tiled) {
      const float *xr = ws.xr.template as<const float>();
      timer.begin(s);
      if (ws.mixed_terms > 0) {
        launch_tiled_mixed_prep(d, k, ws.mixed_terms, cur,
                                ws.hn.template as<const T>(),
                                ws.aop.template as<void>(),
                                ws.bound.template as<float>(), s);
        const bool cand = ws.qmask.bytes() > 0;
        launch_tiled_mixed_assign<I>(n, d, k, ws.mixed_terms, ws.blocks_b, X,
                                     ws.aop.template as<const void>(),
                                     ws.bound.template as<const float>(), labels,
                                     ws.queue.template as<int>(),
                                     ws.queue_counts.template as<int>(),
                                     cand ? ws.qmask.template as<unsigned>() : nullptr,
                                     s, skip, ws.xb.template as<const void>(),
                                     ws.xnb.template as<const float>());
        if (cand)
          launch_tiled_cand_rescore<I>(
              n, d, k, xr, cur, ws.hn.template as<const T>(), labels,
              ws.queue.template as<const int>(),
              ws.queue_counts.template as<const int>(),
              ws.qmask.template as<const unsigned>(), ws.blocks_b, ws.seg_cap,
              ws.cand.template as<void>(), ws.queue_count(), s);
        else
          launch_tiled_rescore<I>(n, d, k, xr, cur, ws.hn.template as<const T>(),
                                  labels, ws.queue.template as<const int>(),
                                  ws.queue_counts.template as<const int>(),
                                  ws.blocks_b, ws.seg_cap, ws.queue_count(), s);
      } else {
        launch_tiled_assign<I>(n, d, k, X, cur, ws.hn.template as<const T>(),
                               labels, s, skip);
      }
      launch_tiled_update<I>(n, d, k, xr, w, labels, ws.scale(),
                             ws.upd.template as<void>(),
                             ws.copies.template as<void>(), s, skip);
      timer.end(s);
      launch_reduce_fixed_partials<T, I>(
          1, d, k, ws.copies.template as<const void>(), ws.scale(),
          ws.nonfinite(), ws.sizes_p(), next, ws.list_p(), ws.n_empty(), k, s, true);
      return;
    }
    if (ws.mixed_terms > 0) {
      launch_mixed_prep(d, k, ws.mixed_terms, cur, ws.hn.template as<const T>(),
                        ws.aop.template as<void>(), ws.bound.template as<float>(),
                        s);
      timer.begin(s);
      mixed_labels_and_partials(ws, X, w, cur, labels, s, skip);
      timer.end(s);
      launch_reduce_fixed_partials<T, I>(
          ws.blocks + ws.blocks_b, d, k, ws.copies.template as<const void>(),
          ws.scale(), ws.nonfinite(), ws.sizes_p(), next, ws.list_p(), ws.n_empty(), k, s, true);
      return;
    }
  }
  if (ws.fast) {
    timer.begin(s);
    launch_fast_fused_fixed<T, I>(n, d, k, X, w, cur,
                                  ws.hn.template as<const T>(), labels,
                                  ws.scale(), ws.nonfinite(),
                                  ws.copies.template as<void>(), ws.blocks, s, skip);
    timer.end(s);
    launch_reduce_fixed_partials<T, I>(
        ws.blocks, d, k, ws.copies.template as<const void>(), ws.scale(),
        ws.nonfinite(), ws.sizes_p(), next, ws.list_p(), ws.n_empty(), k, s, true);
    return;
  }
  if (!ws.store_flush || n == 0) {
    KMD_HIP_CHECK(hipMemsetAsync(ws.copies.template as<T>(), 0,
                                 ws.copies.bytes(), s));
    KMD_HIP_CHECK(hipMemsetAsync(ws.sizes_copies.template as<T>(), 0,
                                 ws.sizes_copies.bytes(), s));
  }
  if (n > 0) {
    timer.begin(s);
    launch_fused_generic<T, I>(n, d, k, ws.blocks, ws.plan, ws.store_flush, X,
                               w, cur, ws.hn.template as<const T>(), labels,
                               ws.copies.template as<T>(),
                               ws.sizes_copies.template as<T>(), s);
    timer.end(s);
  }
  launch_reduce_centroids<T, I>(ws.blocks, d, k,
                                ws.sizes_copies.template as<const T>(),
                                ws.copies.template as<const T>(),
                                ws.sizes_p(), next, ws.list_p(), ws.n_empty(),
                                k, s, true);
}

template <typename T, typename I>
static void driver_impl(int64_t n, int64_t d, int64_t k, const T *X, const T *w,
                        T *init, I *labels, T *res, double tol, bool verbose,
                        int64_t max_iter, int64_t *n_iters_out,
                        double *inertia_out, kmd_print_fn print_fn,
                        void *print_user, hipStream_t s, int precision,
                        DriverStats *stats);

template <typename T, typename I>
void driver_lloyd(int64_t n, int64_t d, int64_t k, const T *X, const T *w,
                  T *init, I *labels, T *res, double tol, bool verbose,
                  int64_t max_iter, int64_t *n_iters_out, double *inertia_out,
                  kmd_print_fn print_fn, void *print_user, hipStream_t s,
                  int precision, DriverStats *stats) {
  TuneGuard g(choose_tuning(n, d, k, precision));
  driver_impl<T, I>(n, d, k, X, w, init, labels, res, tol, verbose, max_iter,
                    n_iters_out, inertia_out, print_fn, print_user, s, precision,
                    stats);
}

// Times the driver's candidate launch plans for this problem and records the
// fastest for its (d, k, precision, log2 n).  Candidates: LDS-resident or
// tiled centroids (where each serves the shape) x persistent blocks per CU
// {occupancy limit, 1, 2, 4}; each runs `iters` Lloyd iterations from the
// given init on scratch copies, timed by the fused step's HIP events.  The
// choice never changes results (every plan computes the same labels and
// exact sums).
template <typename T, typename I>
std::vector<TuneResult> autotune_lloyd(int64_t n, int64_t d, int64_t k,
                                       const T *X, const T *w, const T *init,
                                       int precision, int iters, hipStream_t s) {
  std::vector<TuneResult> out;
  if (n <= 0) return out;
  std::vector<int> paths;
  if (fast_fused_supported<T>(d, k)) paths.push_back(1);
  if (std::is_same<T, float>::value && tiled_supported(d, k)) paths.push_back(2);
  if (paths.empty()) paths.push_back(0);
  TempBuffer c0(sizeof(T) * d * k, s), c1(sizeof(T) * d * k, s),
      lab(sizeof(I) * n, s);
  struct Cand {
    int path, cap, shape;
  };
  std::vector<Cand> cands;
  for (int path : paths) {
    // launch shapes (waves per block x column blocks per wave) of the
    // LDS-resident fixed-point step, then blocks per CU
    const int n_shapes = (path == 1 && std::is_same<T, float>::value) ? 3 : 1;
    for (int shape = 0; shape < n_shapes; ++shape)
      for (int cap : {0, 1, 2, 4}) cands.push_back(Cand{path, cap, shape});
  }
  for (const Cand &cd : cands) {
    {
      const int path = cd.path, cap = cd.cap, shape = cd.shape;
      KMD_HIP_CHECK(hipMemcpyAsync(c0.as<T>(), init, sizeof(T) * d * k,
                                   hipMemcpyDeviceToDevice, s));
      TuneGuard g(TuneChoice{path, cap, shape});
      DriverStats st;
      int64_t it = 0;
      double inertia = 0.0;
      driver_impl<T, I>(n, d, k, X, w, c0.as<T>(), lab.as<I>(), c1.as<T>(), 0.0,
                        false, iters, &it, &inertia, nullptr, nullptr, s,
                        precision, &st);
      TuneResult r;
      r.path = path;
      r.blocks_per_cu = cap;
      r.shape = shape;
      r.ms = st.fused_launches > 0 ? st.fused_kernel_ms / st.fused_launches : 0.0;
      out.push_back(r);
    }
  }
  const TuneResult *best = &out[0];
  for (const TuneResult &r : out)
    if (r.ms > 0.0 && (best->ms <= 0.0 || r.ms < best->ms)) best = &r;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[std::make_tuple(d, k, precision, log2_bucket(n))] =
        TuneChoice{best->path, best->blocks_per_cu, best->shape};
  }
  KMD_HIP_CHECK(hipStreamSynchronize(s));
  return out;
}

template <typename T, typename I>
static void driver_impl(int64_t n, int64_t d, int64_t k, const T *X, const T *w,
                        T *init, I *labels, T *res, double tol, bool verbose,
                        int64_t max_iter, int64_t *n_iters_out,
                        double *inertia_out, kmd_print_fn print_fn,
                        void *print_user, hipStream_t s, int precision,
                        DriverStats *stats) {
  auto ws_ptr = make_workspace<T, I>(n, d, k, X, w, precision, s);
  FusedTimer timer(stats != nullptr);
  auto &ws = *ws_ptr;
  // centroid buffers rotate init -> res -> cbuf: iteration i reads
  // bufs[i % 3] and writes bufs[(i + 1) % 3], so an iteration enqueued ahead
  // never writes the centroids the one before it read
  T *bufs[3] = {init, res, ws.cbuf.template as<T>()};
  const T tol_t = static_cast<T>(tol);
  int64_t it = 0;
  T shift = std::numeric_limits<T>::infinity();
  int64_t rescored = 0;
  PinnedScalars host;
  IterEvents ev;
  // labels of odd iterations go to a second buffer, so the relocation of
  // iteration i still has its labels when iteration i + 1 was enqueued ahead
  // (the epilogue writes the caller's labels)
  TempBuffer lab2(sizeof(I) * (size_t)(n > 0 ? n : 1), s);
  // valid (zero) labels from the start: kernels of an iteration whose main
  // kernel skipped itself may still read them as cluster indices
  KMD_HIP_CHECK(hipMemsetAsync(lab2.as<I>(), 0, lab2.bytes(), s));
  I *labs[2] = {labels, lab2.as<I>()};
  // one iteration: front half, then (unless it found empty clusters) the
  // division, shifts, their sum and the next half-norms; its scalars go to
  // host slot `par`
  auto enqueue_iter = [&](int64_t i) {
    ws.par = (int)(i & 1);
    timer.slot = ws.par;
    T *cur = bufs[i % 3];
    T *next = bufs[(i + 1) % 3];
    // the previous iteration's empty count: when it found empty clusters,
    // this iteration's main kernel returns at once (it will be redone)
    const int *skip =
        i > 0 ? reinterpret_cast<const int *>(ws.scalars.template as<char>() +
                                              32 * (ws.par ^ 1) + 8)
              : nullptr;
    lloyd_front<T, I>(ws, X, w, cur, next, labs[ws.par], s, timer, i > 0, skip);
    launch_finalize_step<T, I>(d, k, ws.n_empty(), ws.sizes_p(), cur, next,
                               ws.shifts.template as<T>(),
                               ws.hn.template as<T>(),
                               ws.fin_part.template as<double>(),
                               ws.fin_ticket.template as<unsigned>(),
                               ws.shift_sum(), s);
    KMD_HIP_CHECK(hipMemcpyAsync(host.slot(ws.par), ws.scalars_p(), 32,
                                 hipMemcpyDeviceToHost, s));
    KMD_HIP_CHECK(hipEventRecord(ev.e[ws.par], s));
  };

  // Iteration it + 1 is enqueued before iteration it's scalars are read
  // (not in verbose mode, which prints between iterations).  It is discarded
  // when the fit stops after iteration it (it wrote only the rotation's third
  // buffer, labels and scratch, all rewritten before use), or when iteration
  // it found empty clusters (then it is enqueued again after the relocation).
  const bool ahead = !verbose;
  if (max_iter > 0) enqueue_iter(0);
  while (it < max_iter && shift > tol_t) {
    bool next_queued = false;
    if (ahead && it + 1 < max_iter) {
      enqueue_iter(it + 1);
      next_queued = true;
    }
    const int p = (int)(it & 1);
    KMD_HIP_CHECK(hipEventSynchronize(ev.e[p]));
    timer.collect(p);
    I ne;
    std::memcpy(&ne, host.slot(p) + 8, sizeof(I));
    const int64_t n_empty = (int64_t)ne;
    if (ws.mixed_terms > 0) {
      int qc;
      std::memcpy(&qc, host.slot(p) + 16, sizeof(int));
      rescored += qc;
    }
    T *cur = bufs[it % 3];
    T *next = bufs[(it + 1) % 3];
    I *lab = labs[p];
    if (verbose) {
      launch_inertia<T, I>(n, d, k, X, w, cur, lab,
                           ws.per_sample.template as<T>(), s);
      double tot = 0.0;
      TempBuffer dres(sizeof(double), s);
      launch_reduce_sum<T>(n, ws.per_sample.template as<const T>(),
                           dres.as<double>(), s);
      KMD_HIP_CHECK(hipMemcpyAsync(&tot, dres.as<double>(), sizeof(double),
                                   hipMemcpyDeviceToHost, s));
      KMD_HIP_CHECK(hipStreamSynchronize(s));
      std::ostringstream ss;
      ss << "Iteration: " << it << " Inertia: " << static_cast<T>(tot);
      if (print_fn) print_fn(ss.str().c_str(), print_user);
    }
    if (n_empty > 0) {
      // the division was skipped; the iteration enqueued ahead (if any) ran
      // on undivided sums and is redone below
      KMD_HIP_CHECK(hipStreamSynchronize(s));
      timer.discard(p ^ 1);  // a skipped launch, when one was enqueued
      ws.par = p;
      // squared distance to the nearest centroid (unit weights), then move
      // the farthest samples into the empty clusters
      launch_inertia<T, I>(n, d, k, X, nullptr, cur, lab,
                           ws.per_sample.template as<T>(), s);
      launch_relocate<T, I>(n, d, k, n_empty, X, w, lab, ws.list_p(),
                            ws.per_sample.template as<const T>(), next,
                            ws.sizes_p(), ws.per_sample.template as<T>(), s);
      // the redone next iteration must not take itself for one to skip
      KMD_HIP_CHECK(hipMemsetAsync(ws.n_empty(), 0, sizeof(I), s));
      launch_finalize_step<T, I>(d, k, static_cast<const I *>(nullptr),
                                 ws.sizes_p(), cur, next,
                                 ws.shifts.template as<T>(),
                                 ws.hn.template as<T>(),
                                 ws.fin_part.template as<double>(),
                                 ws.fin_ticket.template as<unsigned>(),
                                 ws.shift_sum(), s);
      KMD_HIP_CHECK(hipMemcpyAsync(host.slot(p), ws.scalars_p(), 32,
                                   hipMemcpyDeviceToHost, s));
      KMD_HIP_CHECK(hipEventRecord(ev.e[p], s));
      KMD_HIP_CHECK(hipEventSynchronize(ev.e[p]));
      next_queued = false;
    }
    double sh;
    std::memcpy(&sh, host.slot(p), sizeof(double));
    shift = static_cast<T>(sh);
    ++it;
    if (it < max_iter && shift > tol_t && !next_queued) enqueue_iter(it);
  }
  T *cur = bufs[it % 3];
  ws.par = 0;
  auto read_sum = [&]() -> double {
    double v = 0.0;
    KMD_HIP_CHECK(hipMemcpyAsync(&v, ws.shift_sum(), sizeof(double),
                                 hipMemcpyDeviceToHost, s));
    KMD_HIP_CHECK(hipStreamSynchronize(s));
    return v;
  };

  // epilogue: labels and weighted inertia of the final centroids (on the
  // tiled mixed path with its BF16 kernel and exact re-score: the FP32
  // labels)
  launch_half_l2_norm<T>(d, k, cur, ws.hn.template as<T>(), s);
  bool labelled = false;
  if constexpr (std::is_same<T, float>::value) {
    if (ws.tiled && ws.mixed_terms > 0 && n > 0) {
      const float *xr = ws.xr.template as<const float>();
      launch_tiled_mixed_prep(d, k, ws.mixed_terms, cur, ws.hn.template as<const T>(),
                              ws.aop.template as<void>(),
                              ws.bound.template as<float>(), s);
      const bool cand = ws.qmask.bytes() > 0;
      launch_tiled_mixed_assign<I>(n, d, k, ws.mixed_terms, ws.blocks_b, X,
                                   ws.aop.template as<const void>(),
                                   ws.bound.template as<const float>(), labels,
                                   ws.queue.template as<int>(),
                                   ws.queue_counts.template as<int>(),
                                   cand ? ws.qmask.template as<unsigned>() : nullptr,
                                   s, nullptr, ws.xb.template as<const void>(),
                                   ws.xnb.template as<const float>());
      if (cand)
        launch_tiled_cand_rescore<I>(
            n, d, k, xr, cur, ws.hn.template as<const T>(), labels,
            ws.queue.template as<const int>(),
            ws.queue_counts.template as<const int>(),
            ws.qmask.template as<const unsigned>(), ws.blocks_b, ws.seg_cap,
            ws.cand.template as<void>(), ws.queue_count(), s);
      else
        launch_tiled_rescore<I>(n, d, k, xr, cur, ws.hn.template as<const T>(),
                                labels, ws.queue.template as<const int>(),
                                ws.queue_counts.template as<const int>(),
                                ws.blocks_b, ws.seg_cap, ws.queue_count(), s);
      labelled = true;
    } else if (ws.mixed_terms > 0 && n > 0) {
      // the BF16 step and its exact re-score: the FP32 labels (its sums go
      // to the partial slots unread)
      launch_mixed_prep(d, k, ws.mixed_terms, cur, ws.hn.template as<const T>(),
                        ws.aop.template as<void>(), ws.bound.template as<float>(),
                        s);
      mixed_labels_and_partials(ws, X, w, cur, labels, s, nullptr);
      labelled = true;
    }
  }
  if (!labelled)
    launch_assignment<T, I>(n, d, k, 0, X, cur, ws.hn.template as<const T>(),
                            labels, s);
  launch_inertia<T, I>(n, d, k, X, w, cur, labels,
                       ws.per_sample.template as<T>(), s);
  if (cur != res)
    KMD_HIP_CHECK(hipMemcpyAsync(res, cur, sizeof(T) * d * k,
                                 hipMemcpyDeviceToDevice, s));
  launch_reduce_sum<T>(n, ws.per_sample.template as<const T>(),
                       ws.shift_sum(), s);
  *inertia_out = static_cast<double>(static_cast<T>(read_sum()));
  timer.collect_all();
  *n_iters_out = it;
  if (stats) {
    stats->fused_kernel_ms = timer.total_ms;
    stats->fused_launches = timer.launches;
    int path = ws.mixed_terms > 0 ? (ws.mixed_terms == 3 ? 3 : 2)
                                  : (ws.fast ? 1 : 0);
    if (ws.tiled) path += 3;  // 4 tiled FP32, 5 tiled BF16, 6 tiled BF16x3
    if (ws.tiled && ws.mixed_terms == 8) path = 7;  // tiled FP8
    stats->fused_path = path;
    stats->n_rescored = rescored;
  }
}

#define KMD_INST_DRIVER(T, I)                                                 \
  template void driver_lloyd<T, I>(int64_t, int64_t, int64_t, const T *,     \
                                   const T *, T *, I *, T *, double, bool,    \
                                   int64_t, int64_t *, double *,              \
                                   kmd_print_fn, void *, hipStream_t, int,    \
                                   DriverStats *);                            \
  template std::vector<TuneResult> autotune_lloyd<T, I>(                     \
      int64_t, int64_t, int64_t, const T *, const T *, const T *, int, int,   \
      hipStream_t);
KMD_INST_DRIVER(float, int32_t)
KMD_INST_DRIVER(float, int64_t)
KMD_INST_DRIVER(double, int32_t)
KMD_INST_DRIVER(double, int64_t)

void driver_lloyd_single(const BatchedFit &f, hipStream_t s) {
  driver_lloyd<float, int32_t>(f.n, f.d, f.k, f.X, f.w, f.init, f.labels, f.res,
                               f.tol, false, f.max_iter, f.out_iters,
                               f.out_inertia, nullptr, nullptr, s, KMD_PREC_FP32,
                               nullptr);
}

}  // namespace kmd
