// Launch plans and the MFMA fast path of the Lloyd step (lloyd_fast.hip).
#pragma once
#include <vector>

#include "common.hpp"

namespace kmd {

// Timing of the fused-step kernel collected by the driver.
struct DriverStats {
  double fused_kernel_ms = 0.0;  // summed over launches
  int64_t fused_launches = 0;
  int fused_path = 0;   // 0 generic, 1 FP32 MFMA, 2 BF16 (+FP32 re-score), 3 BF16x3
  int64_t n_rescored = 0;       // samples re-scored in FP32 (mixed path)
};

// One autotuner candidate: path (1 LDS-resident, 2 tiled, 0 generic),
// persistent blocks per CU (0 = occupancy limit), fused-step ms per launch.
struct TuneResult {
  int path = 0, blocks_per_cu = 0;
  int shape = 0;   // launch shape of the LDS-resident step (tuned_shape())
  double ms = 0.0;
};
template <typename T, typename I>
std::vector<TuneResult> autotune_lloyd(int64_t n, int64_t d, int64_t k,
                                       const T *X, const T *w, const T *init,
                                       int precision, int iters, hipStream_t s);
// Process-wide override of the tuning choice (path < 0 clears it).
void set_tuning_override(int path, int blocks_per_cu, int shape = 0);

// Launch plan of a persistent fused-step kernel: grid size, dynamic LDS and
// whether centroid sums are privatised in LDS.
struct FusedPlan {
  int blocks = 1;
  size_t lds_bytes = 0;
  bool priv = false;
};

template <typename T>
FusedPlan plan_generic_fused(int64_t n, int64_t d, int64_t k);

template <typename T, typename I>
void launch_fused_generic(int64_t n, int64_t d, int64_t k, int64_t n_copies,
                          const FusedPlan &p, bool store_flush, const T *X,
                          const T *w, const T *C, const T *hn, I *labels,
                          T *copies, T *sizes_copies, hipStream_t s);

// FP64 MFMA path (v_mfma_f64_16x16x4f64; csrc/lloyd_f64.hip): true when the
// centroid operands (and, with update, the FP64 sums) fit LDS, d <= 64.
bool fast64_supported(int64_t d, int64_t k, bool update);
FusedPlan plan_fast64_fused(int64_t n, int64_t d, int64_t k);
template <typename I>
void launch_fast64_fused(int64_t n, int64_t d, int64_t k, int64_t n_copies,
                         const FusedPlan &p, bool store_flush, const double *X,
                         const double *w, const double *C, const double *hn,
                         I *labels, double *copies, double *sizes_copies,
                         hipStream_t s);
template <typename I>
void launch_fast64_assign(int64_t n, int64_t d, int64_t k, const double *X,
                          const double *C, const double *hn, I *labels,
                          hipStream_t s);

// FP32 MFMA path: true when the (d, k) shape is served by it.
template <typename T>
bool fast_fused_supported(int64_t d, int64_t k);

// Number of blocks the MFMA fused kernel launches for n samples (so callers
// that use store_flush can size one private copy per block).
template <typename I>
int fast_fused_blocks(int64_t n, int64_t d, int64_t k);
// k <= 32 (float): the driver's fused step runs the one-sample-per-lane
// kernel (KMD_NO_SMALLK=1 turns it off)
bool fast_smallk_serves(int64_t d, int64_t k);

// The reference's fused step on the MFMA kernel: FP32 block sums in LDS,
// flushed into the caller's private copies with FP32 atomics (no scales, no
// pre-pass, no working memory).
template <typename T, typename I>
void launch_fast_fused(int64_t n, int64_t d, int64_t k, int64_t n_copies,
                       const T *X, const T *w, const T *C, const T *hn,
                       I *labels, T *copies, T *sizes_copies, hipStream_t s);

template <typename T, typename I>
void launch_fast_assign(int64_t n, int64_t d, int64_t k, const T *X,
                        const T *C, const T *hn, I *labels, hipStream_t s);

// Fixed-point scales of the MFMA update: scale[f] = 2^E_f (f < d) and
// scale[d] = 2^E_w such that max|w| max|x_f| 2^E_f < 2^30 (one element fits
// an int32) and n max|w| max|x_f| 2^E_f < 2^61 (the sum fits an int64);
// *nonfinite = 1 when X or w holds inf/NaN, 2 when their range would leave
// typical values coarsely quantised (the kernels then sum in FP32).
// stats_scratch: fixed_point_stats_bytes(d) bytes for the pre-pass's
// statistics, or NULL (allocated per call).  No memset nodes: the statistics
// are cleared by a kernel.
size_t fixed_point_stats_bytes(int64_t d);
void launch_fixed_point_scales(int64_t n, int64_t d, const float *X,
                               const float *w, float *scale, int *nonfinite,
                               hipStream_t s, void *stats_scratch = nullptr);

// MFMA fused step writing one raw partial per block into
// partials[block][d*k + k] (int64 fixed point, or FP32 when *nonfinite).
template <typename T, typename I>
void launch_fast_fused_fixed(int64_t n, int64_t d, int64_t k, const T *X,
                             const T *w, const T *C, const T *hn, I *labels,
                             const float *scale, const int *nonfinite,
                             void *partials, int64_t n_partials,
                             hipStream_t s, const int *skip = nullptr);

// Sums the per-block partials of launch_fast_fused_fixed (exactly, in int64,
// when the fixed-point path ran) into out_cent (d x k) and out_sizes (k),
// then appends empty clusters to empty_list at n_empty[0]++.
template <typename T, typename I>
void launch_reduce_fixed_partials(int64_t n_blocks, int64_t d, int64_t k,
                                  const void *partials, const float *scale,
                                  const int *nonfinite, T *out_sizes,
                                  T *out_cent, I *empty_list, I *n_empty,
                                  int64_t list_len, hipStream_t s,
                                  bool from_zero = false);

}  // namespace kmd

namespace kmd {

// ---- mixed-precision path (lloyd_mixed.hip) --------------------------------
// Bits of the packed centroid index (ceil(log2(k padded to 32))).
int mixed_idx_bits(int64_t k);
// Bytes of the BF16 A-operand array for T split terms (1 or 3).
size_t mixed_aop_bytes(int64_t d, int64_t k, int T);
// Blocks of the BF16 kernel for this shape; 0 when the shape is not served.
// cand: with candidate-tile tracking (launch_mixed_step's qmask); ink: with
// the in-kernel FP32 re-score (launch_mixed_step's C, hn).
template <typename I>
int mixed_blocks(int64_t n, int64_t d, int64_t k, int T, bool cand = false,
                 bool ink = false);
// Whether the candidate tracking fits this shape (k <= 1024, LDS).
bool mixed_cand_fits(int64_t d, int64_t k, int T);
// Whether the in-kernel re-score (FP32 centroid tiles in LDS) fits.
bool mixed_ink_fits(int64_t d, int64_t k, int T);
// Builds the BF16 A operands (centroids + split half-norms) and the error
// bound inputs bound[0] = max ||c||, bound[1] = max hn.
void launch_mixed_prep(int64_t d, int64_t k, int T, const float *C,
                       const float *hn, void *aop, float *bound,
                       hipStream_t s);
// Capacity of each block's segment of the re-score queue.
int64_t mixed_seg_cap(int64_t n, int64_t d, int blocks);
// BF16 assignment + fixed-point update; block b appends its undecided samples
// to queue[b * seg_cap + 0 : queue_counts[b]].  Writes `blocks` partial slots.
// qmask (may be null): per queued sample the bit mask of the 32-centroid
// tiles whose best BF16 score is within the certified bound of the sample's
// best (launch with mixed_blocks(..., cand = true)).  With C and hn (FP32
// centroids and half norms; launch with mixed_blocks(..., ink = true)) the
// undecided samples are re-scored inside the kernel in the FP32 path's
// arithmetic and added to the same partial slots; queue_counts[b] is then
// block b's re-scored count and *rescored_total their sum.
template <typename I>
void launch_mixed_step(int64_t n, int64_t d, int64_t k, int T, int blocks,
                       const float *X, const float *w, const void *aop,
                       const float *bound, const float *scale, I *labels,
                       void *partials, int *queue, int *queue_counts,
                       unsigned *qmask, hipStream_t s, const int *skip = nullptr,
                       const float *C = nullptr, const float *hn = nullptr,
                       int *rescored_total = nullptr);
// int64 fixed-point sums of the queued samples (labels already final), read
// from the sample-major X_r: `blocks` partial slots from `partials`.
template <typename I>
void launch_queue_update(int64_t d, int64_t k, int blocks, int n_segments,
                         int64_t seg_cap, const int *queue, const int *queue_counts,
                         const float *X_r, const float *w, const I *labels,
                         const float *scale, void *partials, hipStream_t s,
                         const int *skip = nullptr);
// FP32 MFMA re-score of the queued samples (the FP32 path's arithmetic):
// reads the n_segments segments of the queue, writes
// fast_rescore_blocks<I>(d, k, n_segments) partial slots from slot0 and the
// number of re-scored samples to *n_rescored.
template <typename I>
int fast_rescore_blocks(int64_t d, int64_t k, int n_segments);
template <typename I>
void launch_fast_rescore(int64_t n, int64_t d, int64_t k, const float *X,
                         const float *w, const float *C, const float *hn,
                         I *labels, const float *scale, const int *nonfinite,
                         void *partials, int64_t slot0, const int *queue,
                         const int *queue_counts, int n_segments,
                         int64_t seg_cap, int *n_rescored, hipStream_t s);

// ---- tiled path (lloyd_tiled.hip): d <= 128, centroids streamed via LDS ----
bool tiled_supported(int64_t d, int64_t k);
// X_r (n, d) = transpose of X_t (d, n)
void launch_transpose(int64_t n, int64_t d, const float *X_t, float *X_r,
                      hipStream_t s);
// FP32 MFMA assignment (labels only); tiles_scratch (tiled_tiles_bytes(d, k)
// bytes, or NULL: allocated per call) holds the prepared centroid tiles
size_t tiled_tiles_bytes(int64_t d, int64_t k);
template <typename I>
void launch_tiled_assign(int64_t n, int64_t d, int64_t k, const float *X_t,
                         const float *C, const float *hn, I *labels,
                         hipStream_t s, const int *skip = nullptr,
                         void *tiles_scratch = nullptr);
// The reference fused step's accumulate-into update from given labels
// (FP32 sums, LDS-privatised per feature group), for the tiled shapes
template <typename I>
void launch_label_update(int64_t n, int64_t d, int64_t k, int64_t n_copies,
                         const float *X_t, const float *w, const I *labels,
                         float *copies, float *sizes_copies, hipStream_t s);
// FP32 re-score of the samples in the queue segments (same arithmetic as
// launch_tiled_assign); samples read from X_r; *n_rescored = queue length
template <typename I>
void launch_tiled_rescore(int64_t n, int64_t d, int64_t k, const float *X_r,
                          const float *C, const float *hn, I *labels,
                          const int *queue, const int *queue_counts,
                          int n_segments, int64_t seg_cap, int *n_rescored,
                          hipStream_t s);
// BF16 (T = 1) / BF16x3 (T = 3) assignment with the certified margin test:
// grid size (0 when not served), queue segment capacity, A-operand bytes
template <typename I>
int tiled_mixed_blocks(int64_t n, int64_t d, int64_t k, int T);
int64_t tiled_mixed_seg_cap(int64_t n, int64_t d, int T, int blocks);
size_t tiled_aop_bytes(int64_t d, int64_t k, int T);
void launch_tiled_mixed_prep(int64_t d, int64_t k, int T, const float *C,
                             const float *hn, void *aop, float *bound,
                             hipStream_t s);
template <typename I>
void launch_tiled_mixed_assign(int64_t n, int64_t d, int64_t k, int T,
                               int blocks, const float *X_t, const void *aop,
                               const float *bound, I *labels, int *queue,
                               int *queue_counts, unsigned *qmask, hipStream_t s,
                               const int *skip = nullptr, const void *xb = nullptr,
                               const float *xnb = nullptr);
// BF16 samples (sample-major, d rounded up to the kernel's bucket) and norm
// bounds for launch_tiled_mixed_assign with T = 1
size_t tiled_xb_bytes(int64_t n, int64_t d);
void launch_tiled_xb_prep(int64_t n, int64_t d, const float *X_t, void *xb,
                          float *xnb, hipStream_t s);
// FP8 (T = 8 in the tiled mixed launchers): e4m3 samples x8 (n bf16_d bytes,
// within tiled_xb_bytes), norm bounds and the fit's sample scale in bound
// (tiled_bound_floats() floats)
size_t tiled_bound_floats();
void launch_tiled_x8_prep(int64_t n, int64_t d, const float *X_t, void *x8,
                          float *xnb, float *bound, hipStream_t s);
// out[i] = e4m3 encode + decode of in[i] with the FP8 kernels' conversion
void launch_f8_roundtrip(int64_t n, const float *in, float *out, hipStream_t s);
// Re-score of the queued samples restricted to their candidate tiles (the
// BF16 kernel's qmask; needs k <= 1024), with the FP32 kernels' arithmetic;
// falls back to the full gather re-score when the (sample, tile) pairs
// overflow the scratch.
bool tiled_cand_supported(int64_t k);
size_t tiled_cand_scratch_bytes(int64_t n, int64_t k);
template <typename I>
void launch_tiled_cand_rescore(int64_t n, int64_t d, int64_t k, const float *X_r,
                               const float *C, const float *hn, I *labels,
                               const int *queue, const int *queue_counts,
                               const unsigned *qmask, int n_segments,
                               int64_t seg_cap, void *scratch, int *n_rescored,
                               hipStream_t s);
// Exact int64 fixed-point sums [d*k + k] of w x per label from X_r
size_t tiled_update_scratch_bytes(int64_t n, int64_t k);
template <typename I>
void launch_tiled_update(int64_t n, int64_t d, int64_t k, const float *X_r,
                         const float *w, const I *labels, const float *scale,
                         void *scratch, void *sums, hipStream_t s,
                         const int *skip = nullptr);

// ---- batched fits (batched.hip) -------------------------------------------
struct BatchedFit {
  int64_t n, d, k;
  const float *X, *w;     // X_t (d, n), weights (n)
  float *init, *res;      // (d, k); init is overwritten as in the driver
  int *labels;            // (n)
  double tol;
  int64_t max_iter;
  int64_t *out_iters;     // host
  double *out_inertia;    // host
};
struct BatchedStats {
  int64_t graph_replays = 0;
  // launch shape of the assignment kernels the call's trial chose (0: the
  // default threads per block, 1: half of them) and the trial's times
  int assign_shape = 0;
  float shape_ms[2] = {0.f, 0.f};
};
// Shapes the batched kernels serve: float data, d <= 64, k <= 256, k <= n.
bool batched_fit_supported(int64_t n, int64_t d, int64_t k);
// One fit through the single-fit FP32 driver (driver_lloyd<float, int32_t>).
void driver_lloyd_single(const BatchedFit &f, hipStream_t s);
// Runs all fits; results equal driver_lloyd<float, int32_t> fit by fit.
// mixed: assignment with the BF16 cross term and the certified FP32
// re-score (labels, and so the fits, identical to the FP32 kernels')
void batched_lloyd(const std::vector<BatchedFit> &fits, int graph_replays,
                   hipStream_t s, BatchedStats *stats, bool mixed = false);
// Device bytes the batch engine holds for one fit while it runs.
int64_t batch_fit_footprint(int64_t n, int64_t d, int64_t k);
// A running set of fits (continuous batching): add() joins a fit (its
// arrays must stay valid until it finishes; inputs with inf/NaN raise a
// ValueError and the fit is not added), step() advances all active fits
// by `iterations` iterations and finishes the ones that stopped (results in
// their out pointers), take_finished() lists them; their workspaces are
// released at the next step.
struct BatchEngine;
BatchEngine *batch_engine_create(hipStream_t user_stream);
void batch_engine_destroy(BatchEngine *e);
int64_t batch_engine_add(BatchEngine *e, const BatchedFit &f);
int64_t batch_engine_step(BatchEngine *e, int iterations, bool use_graph);
std::vector<int64_t> batch_engine_take_finished(BatchEngine *e);
int64_t batch_engine_resident_bytes(const BatchEngine *e);
int64_t batch_engine_launches(const BatchEngine *e);
// Selects the BF16 + certified re-score assignment (true) or the FP32 one.
void batch_engine_set_mixed(BatchEngine *e, bool mixed);

}  // namespace kmd
