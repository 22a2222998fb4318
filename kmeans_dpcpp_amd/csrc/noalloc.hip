// kmd_fused_lloyd_single_step_noalloc: the reference's fused Lloyd step
// (kmd_fused_lloyd_single_step) for callers that cannot allocate device
// memory -- code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc.  The launch's working data (the prepared centroid tiles
// of the tiled shapes; the LDS-resident shapes need none) lives in one
// file-scope device array of kScratchBytes per GPU instead of a per-call
// stream-ordered allocation; the kernels are the ones
// kmd_fused_lloyd_single_step launches.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (two streams, two threads) race on it.  Each call clears what must
// start at zero with a kernel (no memset nodes), so a captured call replays
// correctly.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "lloyd_fast.hpp"

namespace kmd {
namespace {

constexpr size_t kScratchBytes = size_t(1) << 20;   // 1 MiB per GPU

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of this GPU's copy of `scratch`, looked up once per device
// with that device current.  (In host code `scratch` names a host-side
// stand-in; only this address may be handed to kernels.)
void *scratch_address() {
  static std::mutex mu;
  static std::vector<void *> addr;
  int dev = 0;
  KMD_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if ((int)addr.size() <= dev) addr.resize(dev + 1, nullptr);
  if (addr[dev] == nullptr) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      throw RuntimeError(std::string("hipGetSymbolAddress(scratch): ") +
                         hipGetErrorString(e));
    addr[dev] = p;
  }
  return addr[dev];
}

}  // namespace
}  // namespace kmd

using namespace kmd;

namespace {

template <typename F>
int guarded_noalloc(F &&f) {
  try {
    f();
    return KMD_OK;
  } catch (const ValueError &e) {
    set_last_error(e.what());
    return KMD_EVALUE;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return KMD_ERUNTIME;
  } catch (...) {
    set_last_error("unknown error");
    return KMD_ERUNTIME;
  }
}

}  // namespace

extern "C" {

int kmd_fused_lloyd_single_step_noalloc(
    int dtype, int itype, int64_t n_samples, int64_t n_features,
    int64_t n_clusters, int64_t n_copies, int64_t centroids_window_height,
    int64_t work_group_size, const void *X_t, const void *sample_weight,
    const void *centroids_t, const void *centroids_half_l2_norm,
    void *assignments_idx, void *new_centroids_t_private_copies,
    void *cluster_sizes_private_copies, kmd_stream_t stream) {
  return guarded_noalloc([&] {
    if (!((dtype == KMD_F32 || dtype == KMD_F64) &&
          (itype == KMD_I32 || itype == KMD_I64)))
      throw ValueError("Unsupported data types");
    if (!(n_samples >= 0 && n_features >= 1 && n_clusters >= 1))
      throw ValueError("Unexpected array dimensions");
    if (n_copies < 1) throw ValueError("number of private copies must be positive");
    if (!(work_group_size > 0 && centroids_window_height > 0))
      throw ValueError("work_group_size and centroids_window_height must be positive");
    if (n_clusters > INT32_MAX || n_features > INT32_MAX)
      throw ValueError("n_clusters and n_features must fit in 32 bits");
    const void *arrays[] = {X_t, sample_weight, centroids_t, centroids_half_l2_norm,
                            assignments_idx, new_centroids_t_private_copies,
                            cluster_sizes_private_copies};
    const int64_t counts[] = {n_samples * n_features, n_samples,
                              n_features * n_clusters, n_clusters, n_samples,
                              n_copies * n_features * n_clusters, n_copies * n_clusters};
    const char *names[] = {"X_t", "sample_weight", "centroids_t",
                           "centroids_half_l2_norm", "assignments_idx",
                           "new_centroids_t_private_copies",
                           "cluster_sizes_private_copies"};
    for (int i = 0; i < 7; ++i)
      if (counts[i] > 0 && arrays[i] == nullptr)
        throw ValueError(std::string("null pointer for non-empty array ") + names[i]);
    const size_t need = dtype == KMD_F32
                            ? fused_step_scratch_bytes<float>(n_features, n_clusters)
                            : fused_step_scratch_bytes<double>(n_features, n_clusters);
    if (need > kScratchBytes)
      throw ValueError("fused step without allocation: n_features = " +
                       std::to_string(n_features) + ", n_clusters = " +
                       std::to_string(n_clusters) + " need " + std::to_string(need) +
                       " bytes of scratch, more than the " +
                       std::to_string(kScratchBytes) + " available");
    if (n_samples == 0) return;
    void *buf = scratch_address();
    const hipStream_t s = static_cast<hipStream_t>(stream);
    auto run = [&](auto *tp, auto *ip) {
      using T = std::remove_pointer_t<decltype(tp)>;
      using I = std::remove_pointer_t<decltype(ip)>;
      launch_fused_step_copies<T, I>(
          n_samples, n_features, n_clusters, n_copies, centroids_window_height,
          (const T *)X_t, (const T *)sample_weight, (const T *)centroids_t,
          (const T *)centroids_half_l2_norm, (I *)assignments_idx,
          (T *)new_centroids_t_private_copies, (T *)cluster_sizes_private_copies, s,
          buf);
    };
    if (dtype == KMD_F32 && itype == KMD_I32)
      run((float *)nullptr, (int32_t *)nullptr);
    else if (dtype == KMD_F32)
      run((float *)nullptr, (int64_t *)nullptr);
    else if (itype == KMD_I32)
      run((double *)nullptr, (int32_t *)nullptr);
    else
      run((double *)nullptr, (int64_t *)nullptr);
  });
}

}  // extern "C"
