// pybind11 boundary `_kmeans_dpcpp_amd`: the 14 entry points of the
// reference module `_kmeans_dpcpp` (python_api/_kmeans_lloyd.cpp:1277-1449)
// with the same names, positional order, defaults and ValueError messages.
//
// Arrays arrive through the DLPack protocol (`__dlpack__` /
// `__dlpack_device__`) and must be ROCm device arrays in compact row-major
// layout with dtype float32/float64 (data) and int32/int64 (labels, indices).
// The `sycl_queue` slot takes a stream-like: None (null stream), an int
// (hipStream_t value), an object with a `cuda_stream` attribute
// (torch.cuda.Stream) or a kmeans_dpcpp_amd Stream.  `depends` is a list of
// Event objects returned by earlier calls.  Asynchronous entry points return
// (keepalive_event, compute_event); the keepalive event holds the argument
// capsules until it has been waited on.
//
// This file has no HIP code: it calls libkmeans_dpcpp_amd.so through its C
// ABI (include/kmeans_dpcpp_amd.h).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/kmeans_dpcpp_amd.h"
#include "dlpack_abi.h"

namespace py = pybind11;
using namespace kmd_dlpack;

namespace {

void check_rc(int rc) {
  if (rc == KMD_OK) return;
  std::string msg = kmd_last_error();
  if (rc == KMD_EVALUE) throw py::value_error(msg);
  throw std::runtime_error(msg);
}

// ---- events ----------------------------------------------------------------

struct EventHandle {
  kmd_event_t ev = nullptr;
  ~EventHandle() {
    if (ev) kmd_event_destroy(ev);
  }
};

class Event {
 public:
  Event() = default;
  Event(std::shared_ptr<EventHandle> h, py::list keep)
      : h_(std::move(h)), keep_(std::move(keep)) {}
  void wait() {
    if (h_ && h_->ev) {
      int rc;
      {
        py::gil_scoped_release nogil;
        rc = kmd_event_synchronize(h_->ev);
      }
      check_rc(rc);
    }
    keep_ = py::list();
  }
  bool is_complete() const {
    if (!h_ || !h_->ev) return true;
    int done = 0;
    check_rc(kmd_event_query(h_->ev, &done));
    return done != 0;
  }
  float elapsed_ms(const Event &stop) const {
    float ms = 0.f;
    if (!h_ || !stop.h_) throw py::value_error("event was never recorded");
    check_rc(kmd_event_elapsed_ms(h_->ev, stop.h_->ev, &ms));
    return ms;
  }
  kmd_event_t raw() const { return h_ ? h_->ev : nullptr; }

 private:
  std::shared_ptr<EventHandle> h_;
  py::list keep_;
};

kmd_stream_t resolve_stream(const py::object &q) {
  if (q.is_none()) return nullptr;
  if (py::isinstance<py::int_>(q))
    return reinterpret_cast<kmd_stream_t>(q.cast<intptr_t>());
  if (py::hasattr(q, "cuda_stream"))
    return reinterpret_cast<kmd_stream_t>(q.attr("cuda_stream").cast<intptr_t>());
  if (py::hasattr(q, "handle"))
    return reinterpret_cast<kmd_stream_t>(q.attr("handle").cast<intptr_t>());
  throw py::value_error(
      "sycl_queue must be None, an int stream handle or a stream object");
}

void wait_depends(kmd_stream_t s, const py::list &depends) {
  for (auto item : depends) {
    if (item.is_none()) continue;
    Event &e = item.cast<Event &>();
    if (e.raw()) check_rc(kmd_stream_wait_event(s, e.raw()));
  }
}

std::pair<Event, Event> record(kmd_stream_t s, py::list keep) {
  auto h = std::make_shared<EventHandle>();
  check_rc(kmd_event_create(&h->ev));
  check_rc(kmd_event_record(h->ev, s));
  return {Event(h, std::move(keep)), Event(h, py::list())};
}

// ---- DLPack arrays ---------------------------------------------------------

struct Arr {
  py::object capsule;
  void *data = nullptr;
  int ndim = 0;
  std::vector<int64_t> shape;
  int64_t size = 1;
  bool c_contig = true;
  uint8_t code = 0, bits = 0;
  int device = -1;

  int64_t dim(int i) const { return shape.at(i); }
  bool is_float() const { return code == kDLFloat && (bits == 32 || bits == 64); }
  bool is_int() const { return code == kDLInt && (bits == 32 || bits == 64); }
  int dtype() const { return bits == 32 ? KMD_F32 : KMD_F64; }
  int itype() const { return bits == 32 ? KMD_I32 : KMD_I64; }
  bool same_type(const Arr &o) const { return code == o.code && bits == o.bits; }
};

Arr to_arr(const py::handle &obj, kmd_stream_t s, const char *name) {
  if (!py::hasattr(obj, "__dlpack__") || !py::hasattr(obj, "__dlpack_device__"))
    throw py::value_error(std::string("argument `") + name +
                          "` does not support the DLPack protocol");
  py::tuple dev = obj.attr("__dlpack_device__")();
  const int dev_type = dev[0].cast<int>();
  if (dev_type != kDLROCM)
    throw py::value_error(std::string("argument `") + name +
                          "` must be a ROCm device array");
  Arr a;
  a.capsule = obj.attr("__dlpack__")(py::arg("stream") =
                                         reinterpret_cast<intptr_t>(s));
  auto *mt = static_cast<ManagedTensor *>(
      PyCapsule_GetPointer(a.capsule.ptr(), "dltensor"));
  if (mt == nullptr) throw py::error_already_set();
  const Tensor &t = mt->dl_tensor;
  a.data = static_cast<char *>(t.data) + t.byte_offset;
  a.ndim = t.ndim;
  a.shape.assign(t.shape, t.shape + t.ndim);
  a.code = t.dtype.code;
  a.bits = t.dtype.bits;
  a.device = t.device.device_id;
  for (int i = 0; i < t.ndim; ++i) a.size *= t.shape[i];
  if (t.strides != nullptr && a.size > 1) {
    int64_t expect = 1;
    for (int i = t.ndim - 1; i >= 0; --i) {
      if (t.shape[i] != 1 && t.strides[i] != expect) a.c_contig = false;
      expect *= t.shape[i];
    }
  }
  if (t.dtype.lanes != 1) a.c_contig = false;
  return a;
}

bool all_contig(const std::vector<const Arr *> &xs) {
  for (auto *x : xs)
    if (!x->c_contig) return false;
  return true;
}
bool same_device(const std::vector<const Arr *> &xs) {
  int d = -2;
  for (auto *x : xs) {
    if (d == -2) d = x->device;
    if (x->device != d) return false;
  }
  return true;
}
py::list keep_of(std::initializer_list<const Arr *> xs) {
  py::list l;
  for (auto *x : xs) l.append(x->capsule);
  return l;
}
void use_device(const Arr &a) { check_rc(kmd_set_device(a.device)); }

const char *kQueueMsg =
    "Execution queue is not compatible with allocation queues";

// ---- entry points (python_api/_kmeans_lloyd.cpp) ---------------------------

std::pair<Event, Event> py_broadcast_divide(py::object X_, py::object y_,
                                            size_t work_group_size,
                                            py::object q, py::list depends) {
  kmd_stream_t s = resolve_stream(q);
  Arr X = to_arr(X_, s, "divident"), y = to_arr(y_, s, "divisor");
  if (X.ndim != 2 || y.ndim != 1 || !all_contig({&X, &y}))
    throw py::value_error(
        "Arguments must be a matrix and a vector with C-contiguous layout");
  if (y.dim(0) != X.dim(1))
    throw py::value_error(
        "Array dimensions of arguments are not consistent, X.shape[1] != "
        "y.shape[0]");
  if (!same_device({&X, &y})) throw py::value_error(kQueueMsg);
  if (!X.same_type(y))
    throw py::value_error("Arguments must have the same elemental data types.");
  if (!y.is_float())
    throw py::value_error(
        "Unsupported elemental data type. Expecting single or double precision "
        "floating point numbers");
  use_device(X);
  wait_depends(s, depends);
  check_rc(kmd_broadcast_divide(y.dtype(), X.dim(0), X.dim(1),
                                (int64_t)work_group_size, X.data, y.data, s));
  return record(s, keep_of({&X, &y}));
}

std::pair<Event, Event> py_half_l2_norm_squared(py::object X_, py::object y_,
                                                size_t work_group_size,
                                                py::object q, py::list depends) {
  kmd_stream_t s = resolve_stream(q);
  Arr X = to_arr(X_, s, "centroids"),
      y = to_arr(y_, s, "centroids_half_l2_norm_squared");
  if (X.ndim != 2 || y.ndim != 1 || !all_contig({&X, &y}))
    throw py::value_error(
        "Arguments must be a matrix and a vector with C-contiguous layout");
  if (y.dim(0) != X.dim(1))
    throw py::value_error(
        "Array dimensions of arguments are not consistent, X.shape[1] != "
        "y.shape[0]");
  if (!same_device({&X, &y})) throw py::value_error(kQueueMsg);
  if (!X.same_type(y))
    throw py::value_error("Arguments must have the same elemental data types.");
  if (!y.is_float())
    throw py::value_error(
        "Unsupported elemental data type. Expecting single or double precision "
        "floating point numbers");
  use_device(X);
  wait_depends(s, depends);
  check_rc(kmd_half_l2_norm_squared(y.dtype(), X.dim(0), X.dim(1),
                                    (int64_t)work_group_size, X.data, y.data,
                                    s));
  return record(s, keep_of({&X, &y}));
}

std::pair<Event, Event> py_reduce_centroids_data(
    py::object sizes_copies_, py::object cent_copies_, py::object out_sizes_,
    py::object out_cent_, py::object list_, py::object n_empty_,
    size_t work_group_size, py::object q, py::list depends) {
  kmd_stream_t s = resolve_stream(q);
  Arr sc = to_arr(sizes_copies_, s, "cluster_sizes_private_copies");
  Arr cc = to_arr(cent_copies_, s, "centroids_t_private_copies");
  Arr os = to_arr(out_sizes_, s, "out_cluster_sizes");
  Arr oc = to_arr(out_cent_, s, "out_centroids_t");
  Arr ol = to_arr(list_, s, "out_empty_clusters_list");
  Arr ne = to_arr(n_empty_, s, "out_n_empty_clusters");
  if (sc.ndim != 2 || cc.ndim != 3 || os.ndim != 1 || oc.ndim != 2 ||
      ol.ndim != 1 || ne.ndim != 1)
    throw py::value_error("Array dimensions of inputs are not consistent");
  const int64_t n_copies = sc.dim(0), n_clusters = sc.dim(1),
                n_features = cc.dim(1);
  if (n_copies != cc.dim(0) || n_clusters != cc.dim(2) ||
      n_clusters != os.dim(0) || n_clusters != oc.dim(1) ||
      n_features != oc.dim(0) || n_clusters != ol.dim(0))
    throw py::value_error("Dimensions mismatch");
  if (!sc.same_type(cc) || !sc.same_type(os) || !sc.same_type(oc) ||
      !ne.same_type(ol))
    throw py::value_error("Array element data types must be consisten");
  if (!all_contig({&sc, &cc, &os, &oc, &ol, &ne}))
    throw py::value_error("All array arguments must be C-contiguous");
  if (!same_device({&sc, &cc, &os, &oc, &ol, &ne}))
    throw py::value_error(kQueueMsg);
  if (!sc.is_float() || !ne.is_int())
    throw py::value_error("Unsupported data types");
  if (ne.size < 1) throw py::value_error("Dimensions mismatch");
  use_device(sc);
  wait_depends(s, depends);
  check_rc(kmd_reduce_centroids_data(sc.dtype(), ne.itype(), n_copies,
                                     n_features, n_clusters,
                                     (int64_t)work_group_size, sc.data,
                                     cc.data, os.data, oc.data, ol.data,
                                     ne.data, s));
  return record(s, keep_of({&sc, &cc, &os, &oc, &ol, &ne}));
}

std::pair<Event, Event> py_compute_threshold(py::object data_, size_t topk,
                                             py::object thr_, py::object q,
                                             py::list depends) {
  kmd_stream_t s = resolve_stream(q);
  Arr data = to_arr(data_, s, "data"), thr = to_arr(thr_, s, "threshold");
  if (data.ndim != 1 || !all_contig({&data, &thr}))
    throw py::value_error("Argument data must be a C-contiguous vector");
  if (thr.size != 1)
    throw py::value_error("Argument threshold must be 1-element array");
  if (!data.same_type(thr))
    throw py::value_error("Data types of arguments must be the same");
  if (!same_device({&data, &thr})) throw py::value_error(kQueueMsg);
  if (topk == 0) return {Event(), Event()};
  if (!data.is_float())
    throw py::value_error(
        "Unsupported elemental data type. Expect single- or double- "
        "floating-point types.");
  use_device(data);
  wait_depends(s, depends);
  check_rc(kmd_compute_threshold(data.dtype(), data.dim(0), data.data,
                                 (int64_t)topk, thr.data, s));
  return record(s, keep_of({&data, &thr}));
}

std::pair<Event, Event> py_select_samples_far_from_centroid(
    size_t n_selected, py::object dist_, py::object thr_, py::object sel_,
    py::object ngt_, py::object neq_, size_t work_group_size, py::object q,
    py::list depends) {
  if (n_selected == 0)
    throw py::value_error("Argument `n_selected` must be positive");
  kmd_stream_t s = resolve_stream(q);
  Arr dist = to_arr(dist_, s, "distance_to_centroid");
  Arr thr = to_arr(thr_, s, "threshold");
  Arr sel = to_arr(sel_, s, "selected_samples_idx");
  Arr ngt = to_arr(ngt_, s, "n_selected_gt_threshold");
  Arr neq = to_arr(neq_, s, "n_selected_eq_threshold");
  if (dist.ndim != 1 || sel.ndim != 1 || thr.size != 1 || ngt.size != 1 ||
      neq.size != 1)
    throw py::value_error("Array dimensionalities are not consistent");
  const int64_t n_samples = dist.dim(0);
  if (n_samples < (int64_t)n_selected)
    throw py::value_error("Argument `n_select` is too large");
  if (n_samples != sel.dim(0))
    throw py::value_error(
        "Vector `selected_samples_idx` must have size `n_samples`");
  if (!all_contig({&dist, &sel}))
    throw py::value_error("Arrays must be C-contiguous");
  if (!dist.same_type(thr))
    throw py::value_error("Input arrays must have the same elemental data type");
  if (!sel.same_type(ngt) || !sel.same_type(neq))
    throw py::value_error("Output array must have the same elemental data type");
  if (!same_device({&dist, &thr, &sel, &ngt, &neq}))
    throw py::value_error(kQueueMsg);
  if (!dist.is_float() || !sel.is_int())
    throw py::value_error("Unsupported data types");
  use_device(dist);
  wait_depends(s, depends);
  check_rc(kmd_select_samples_far_from_centroid(
      dist.dtype(), sel.itype(), (int64_t)n_selected, n_samples,
      (int64_t)work_group_size, dist.data, thr.data, sel.data, ngt.data,
      neq.data, s));
  return record(s, keep_of({&dist, &thr, &sel, &ngt, &neq}));
}

std::pair<Event, Event> py_relocate_empty_clusters(
    size_t n_empty, py::object X_, py::object w_, py::object lab_,
    py::object list_, py::object sq_, py::object C_, py::object sizes_,
    py::object inertia_, size_t work_group_size, py::object q,
    py::list depends) {
  kmd_stream_t s = resolve_stream(q);
  Arr X = to_arr(X_, s, "X_t"), w = to_arr(w_, s, "sample_weights");
  Arr lab = to_arr(lab_, s, "assignment_id");
  Arr lst = to_arr(list_, s, "empty_clusters_list");
  Arr sq = to_arr(sq_, s, "sq_dist_to_nearest_centroid");
  Arr C = to_arr(C_, s, "centroid_t"), sz = to_arr(sizes_, s, "cluster_sizes");
  Arr in = to_arr(inertia_, s, "per_sample_inertia");
  if (X.ndim != 2 || w.ndim != 1 || lab.ndim != 1 || lst.ndim != 1 ||
      sq.ndim != 1 || C.ndim != 2 || sz.ndim != 1 || in.ndim != 1)
    throw py::value_error("Arguments have inconsistent array dimensionality.");
  if (!all_contig({&X, &w, &lab, &lst, &sq, &C, &sz, &in}))
    throw py::value_error("Inputs must be C-contiguous");
  const int64_t n_samples = X.dim(1), n_features = X.dim(0),
                n_clusters = lst.dim(0);
  if (n_samples != w.dim(0) || n_samples != lab.dim(0) ||
      n_samples != sq.dim(0) || n_clusters != C.dim(1) ||
      n_features != C.dim(0) || n_clusters != sz.dim(0) ||
      n_samples != in.dim(0))
    throw py::value_error("Input dimensions are inconsistent");
  if (n_empty == 0)
    throw py::value_error("n_empty_clusters must be non-zero.");
  if (!same_device({&X, &w, &lab, &lst, &sq, &C, &sz, &in}))
    throw py::value_error(
        "Execution queue is not compatible with allocation queues.");
  if (!X.same_type(w) || !X.same_type(sq) || !X.same_type(C) ||
      !X.same_type(sz) || !X.same_type(in) || !lab.same_type(lst))
    throw py::value_error("Inconsistent array elemental data types");
  if (!X.is_float() || !lab.is_int())
    throw py::value_error("Unsupported data types");
  use_device(X);
  wait_depends(s, depends);
  check_rc(kmd_relocate_empty_clusters(
      X.dtype(), lab.itype(), n_samples, n_features, n_clusters,
      (int64_t)n_empty, (int64_t)work_group_size, X.data, w.data, lab.data,
      lst.data, sq.data, C.data, sz.data, in.data, s));
  return record(s, keep_of({&X, &w, &lab, &lst, &sq, &C, &sz, &in}));
}

std::pair<Event, Event> py_compute_centroid_shifts_squared(
    py::object old_, py::object new_, py::object out_, py::object q,
    py::list depends) {
  kmd_stream_t s = resolve_stream(q);
  Arr a = to_arr(old_, s, "centroid_t"), b = to_arr(new_, s, "new_centroid_t");
  Arr o = to_arr(out_, s, "out_centroid_shifts");
  if (b.ndim != 2 || a.ndim != 2 || o.ndim != 1)
    throw py::value_error("Input dimensionalities are not consistent.");
  const int64_t n_features = a.dim(0), n_clusters = a.dim(1);
  if (n_features != b.dim(0) || n_clusters != b.dim(1) || n_clusters != o.dim(0))
    throw py::value_error("Array dimensions are not consistent.");
  if (!all_contig({&a, &b, &o}))
    throw py::value_error("Arguments must be C-contiguous arrays");
  if (!same_device({&a, &b, &o})) throw py::value_error(kQueueMsg);
  if (!a.same_type(b) || !a.same_type(o))
    throw py::value_error(
        "All array arguments must have the same elemental data types");
  if (!a.is_float()) throw py::value_error("Unsupported elemental data type.");
  use_device(a);
  wait_depends(s, depends);
  check_rc(kmd_compute_centroid_shifts_squared(a.dtype(), n_features,
                                               n_clusters, a.data, b.data,
                                               o.data, s));
  return record(s, keep_of({&a, &b, &o}));
}

std::pair<Event, Event> py_compute_distances(py::object X_, py::object C_,
                                             py::object D_,
                                             size_t work_group_size,
                                             size_t centroids_window_height,
                                             py::object q, py::list depends) {
  kmd_stream_t s = resolve_stream(q);
  Arr X = to_arr(X_, s, "X_t"), C = to_arr(C_, s, "centroid_t");
  Arr D = to_arr(D_, s, "euclidean_distances_t");
  if (X.ndim != 2 || C.ndim != 2 || D.ndim != 2)
    throw py::value_error("Input arrays must have dimensionality 2.");
  if (!all_contig({&X, &C, &D}))
    throw py::value_error("Input arrays must be C-contiguous.");
  const int64_t n_features = X.dim(0), n_samples = X.dim(1),
                n_clusters = D.dim(0);
  if (n_features != C.dim(0) || n_clusters != C.dim(1) ||
      n_samples != D.dim(1))
    throw py::value_error("Input array dimensions are not consistant");
  if (!same_device({&X, &C, &D})) throw py::value_error(kQueueMsg);
  if (!X.same_type(C) || !X.same_type(D))
    throw py::value_error("Arrays must have the same elemental data types");
  if (!X.is_float()) throw py::value_error("Unsupported elemental data type");
  use_device(X);
  wait_depends(s, depends);
  check_rc(kmd_compute_centroid_to_sample_distances(
      X.dtype(), n_samples, n_features, n_clusters, (int64_t)work_group_size,
      (int64_t)centroids_window_height, X.data, C.data, D.data, s));
  return record(s, keep_of({&X, &C, &D}));
}

std::pair<Event, Event> py_assignment(py::object X_, py::object C_,
                                      py::object hn_, py::object lab_,
                                      size_t centroids_window_height,
                                      size_t work_group_size, py::object q,
                                      py::list depends) {
  kmd_stream_t s = resolve_stream(q);
  Arr X = to_arr(X_, s, "X_t"), C = to_arr(C_, s, "centroids_t");
  Arr hn = to_arr(hn_, s, "centroids_half_l2_norm");
  Arr lab = to_arr(lab_, s, "assignment_id");
  if (X.ndim != 2 || C.ndim != 2 || hn.ndim != 1 || lab.ndim != 1)
    throw py::value_error("Inputs have unexpected dimensionality.");
  if (!all_contig({&X, &C, &hn, &lab}))
    throw py::value_error("Inputs must be C-contiguous arrays.");
  const int64_t n_features = X.dim(0), n_samples = X.dim(1),
                n_clusters = hn.dim(0);
  if (n_features != C.dim(0) || n_clusters != C.dim(1) ||
      n_samples != lab.dim(0))
    throw py::value_error("Inputs have inconsistent dimensions.");
  if (!same_device({&X, &C, &hn, &lab}))
    throw py::value_error(
        "Execution queue is incompatible with allocation queues.");
  if (!X.same_type(C) || !X.same_type(hn))
    throw py::value_error("Arrays have inconsistent elemental data types");
  if (!X.is_float() || !lab.is_int())
    throw py::value_error("Unsupported array elemental data type");
  use_device(X);
  wait_depends(s, depends);
  check_rc(kmd_assignment(X.dtype(), lab.itype(), n_samples, n_features,
                          n_clusters, (int64_t)centroids_window_height,
                          (int64_t)work_group_size, X.data, C.data, hn.data,
                          lab.data, s));
  return record(s, keep_of({&X, &C, &hn, &lab}));
}

std::pair<Event, Event> py_compute_inertia(py::object X_, py::object w_,
                                           py::object C_, py::object lab_,
                                           py::object out_,
                                           size_t work_group_size,
                                           py::object q, py::list depends) {
  kmd_stream_t s = resolve_stream(q);
  Arr X = to_arr(X_, s, "X_t"), w = to_arr(w_, s, "sample_weight");
  Arr C = to_arr(C_, s, "centroid_t"), lab = to_arr(lab_, s, "assignment_id");
  Arr out = to_arr(out_, s, "out_per_sample_inertia");
  if (X.ndim != 2 || w.ndim != 1 || C.ndim != 2 || lab.ndim != 1 ||
      out.ndim != 1)
    throw py::value_error("Input array dimensionality is not consistent");
  if (!all_contig({&X, &w, &C, &lab, &out}))
    throw py::value_error("All input arrays must be C-contiguous");
  if (!same_device({&X, &w, &C, &lab, &out})) throw py::value_error(kQueueMsg);
  const int64_t n_features = X.dim(0), n_samples = X.dim(1),
                n_clusters = C.dim(1);
  if (n_features != C.dim(0) || n_samples != w.dim(0) ||
      n_samples != lab.dim(0) || n_samples != out.dim(0))
    throw py::value_error("Array dimensions are not consistent");
  if (!X.same_type(w) || !X.same_type(C) || !X.same_type(out))
    throw py::value_error("Array elemental data types are not consistent");
  if (!X.is_float() || !lab.is_int())
    throw py::value_error("Unsupported array elemental data type");
  use_device(X);
  wait_depends(s, depends);
  check_rc(kmd_compute_inertia(X.dtype(), lab.itype(), n_samples, n_features,
                               n_clusters, (int64_t)work_group_size, X.data,
                               w.data, C.data, lab.data, out.data, s));
  return record(s, keep_of({&X, &w, &C, &lab, &out}));
}

py::object py_reduce_vector_blocking(py::object data_, py::object q,
                                     py::list depends) {
  kmd_stream_t s = resolve_stream(q);
  Arr data = to_arr(data_, s, "data");
  if (data.ndim != 1 || !data.c_contig)
    throw py::value_error("Expecting 1D contiguous vector.");
  if (!data.is_float()) throw py::value_error("Unsupported data type");
  use_device(data);
  wait_depends(s, depends);
  double res = 0.0;
  int rc;
  {
    py::gil_scoped_release nogil;
    rc = kmd_reduce_vector_blocking(data.dtype(), data.dim(0), data.data, &res,
                                    s);
  }
  check_rc(rc);
  return py::float_(res);
}

std::pair<Event, Event> py_fused_lloyd_single_step(
    py::object X_, py::object w_, py::object C_, py::object hn_,
    py::object lab_, py::object copies_, py::object sizes_copies_,
    size_t centroids_window_height, size_t work_group_size, py::object q,
    py::list depends) {
  kmd_stream_t s = resolve_stream(q);
  Arr X = to_arr(X_, s, "X_t"), w = to_arr(w_, s, "sample_weight");
  Arr C = to_arr(C_, s, "centroids_t");
  Arr hn = to_arr(hn_, s, "centroids_half_l2_norm");
  Arr lab = to_arr(lab_, s, "assignments_idx");
  Arr cp = to_arr(copies_, s, "new_centroids_t_private_copies");
  Arr scp = to_arr(sizes_copies_, s, "cluster_sizes_private_copies");
  if (X.ndim != 2 || w.ndim != 1 || C.ndim != 2 || hn.ndim != 1 ||
      lab.ndim != 1 || cp.ndim != 3 || scp.ndim != 2)
    throw py::value_error("Unexpected input array dimensionalities.");
  if (!all_contig({&X, &w, &C, &hn, &lab, &cp, &scp}))
    throw py::value_error("All arrays must be C-contiguous");
  const int64_t n_features = X.dim(0), n_samples = X.dim(1),
                n_clusters = hn.dim(0), n_copies = cp.dim(0);
  if (n_features != C.dim(0) || n_clusters != C.dim(1) ||
      n_samples != w.dim(0) || n_samples != lab.dim(0) ||
      n_features != cp.dim(1) || n_clusters != cp.dim(2) ||
      n_copies != scp.dim(0) || n_clusters != scp.dim(1))
    throw py::value_error("Unexpected array dimensions");
  if (!same_device({&X, &w, &C, &hn, &lab, &cp, &scp}))
    throw py::value_error(
        "Execution queue is not compatible with allocation queues.");
  if (!X.same_type(w) || !X.same_type(C) || !X.same_type(hn) ||
      !X.same_type(cp) || !X.same_type(scp))
    throw py::value_error("Array arguments have different elemental data types");
  if (!X.is_float() || !lab.is_int())
    throw py::value_error("Unsupported array elemental data types.");
  use_device(X);
  wait_depends(s, depends);
  check_rc(kmd_fused_lloyd_single_step(
      X.dtype(), lab.itype(), n_samples, n_features, n_clusters, n_copies,
      (int64_t)centroids_window_height, (int64_t)work_group_size, X.data,
      w.data, C.data, hn.data, lab.data, cp.data, scp.data, s));
  return record(s, keep_of({&X, &w, &C, &hn, &lab, &cp, &scp}));
}

size_t py_compute_number_of_private_copies(py::object arr_, size_t n_samples,
                                           size_t n_features,
                                           size_t n_clusters, double occupancy,
                                           size_t work_group_size) {
  Arr a = to_arr(arr_, nullptr, "array");
  if (occupancy <= 0.0 || occupancy >= 1.0)
    throw py::value_error("Expecting a fraction strictly between 0 and 1");
  if (!a.is_float()) throw py::value_error("Unsupported elemental data type");
  int64_t out = 0;
  check_rc(kmd_compute_number_of_private_copies(
      a.dtype(), a.device, (int64_t)n_samples, (int64_t)n_features,
      (int64_t)n_clusters, occupancy, (int64_t)work_group_size, &out));
  return (size_t)out;
}

void print_to_python(const char *msg, void *) {
  py::gil_scoped_acquire gil;
  py::print(msg);
}

std::pair<size_t, py::array> py_kmeans_lloyd_driver(
    py::object X_, py::object w_, py::object init_, py::object lab_,
    py::object res_, double tol, bool verbose, size_t max_iter,
    size_t centroids_window_height, size_t work_group_size, double occupancy,
    py::object q, py::list depends) {
  kmd_stream_t s = resolve_stream(q);
  Arr X = to_arr(X_, s, "X_t"), w = to_arr(w_, s, "sample_weight");
  Arr init = to_arr(init_, s, "init_centroid_t");
  Arr lab = to_arr(lab_, s, "assignments_id");
  Arr res = to_arr(res_, s, "res_centroids_t");
  if (X.ndim != 2 || w.ndim != 1 || init.ndim != 2 || res.ndim != 2 ||
      lab.ndim != 1)
    throw py::value_error("Unsupported array dimensionalities");
  if (!all_contig({&X, &w, &init, &lab, &res}))
    throw py::value_error("All input arrays must be C-contiguous");
  if (!same_device({&X, &w, &init, &lab, &res}))
    throw py::value_error(kQueueMsg);
  const int64_t n_features = X.dim(0), n_samples = X.dim(1),
                n_clusters = init.dim(1);
  if (n_features != init.dim(0) || n_features != res.dim(0) ||
      n_clusters != res.dim(1) || n_samples != w.dim(0) ||
      n_samples != lab.dim(0))
    throw py::value_error("Array dimensions are not consistent");
  if (!X.same_type(w) || !X.same_type(init) || !X.same_type(res))
    throw py::value_error(
        "Sample coordinates, weights and centroids must have the same "
        "elemental data types");
  if (occupancy <= 0.0 || occupancy >= 1.0)
    throw py::value_error(
        "Fraction `centroids_private_copies_max_cache_occupancy` is out of "
        "bounds");
  if (tol < 0.0) throw py::value_error("Tolerance must be non-negative");
  if (!X.is_float() || !lab.is_int())
    throw py::value_error("Unsupport elemental data type");
  use_device(X);
  wait_depends(s, depends);
  int64_t n_iters = 0;
  double inertia = 0.0;
  int rc;
  {
    py::gil_scoped_release nogil;
    rc = kmd_kmeans_lloyd_driver(
        X.dtype(), lab.itype(), n_samples, n_features, n_clusters, X.data,
        w.data, init.data, lab.data, res.data, tol, verbose ? 1 : 0,
        (int64_t)max_iter, (int64_t)centroids_window_height,
        (int64_t)work_group_size, occupancy, &n_iters, &inertia,
        verbose ? &print_to_python : nullptr, nullptr, s);
  }
  check_rc(rc);
  py::array out;
  if (X.dtype() == KMD_F32) {
    py::array_t<float> a(1);
    a.mutable_at(0) = (float)inertia;
    out = a;
  } else {
    py::array_t<double> a(1);
    a.mutable_at(0) = inertia;
    out = a;
  }
  return {(size_t)n_iters, out};
}

std::pair<Event, Event> py_iterative_merge_sort(py::object in_, py::object out_,
                                                py::object q,
                                                py::list depends) {
  kmd_stream_t s = resolve_stream(q);
  Arr in = to_arr(in_, s, "data"), out = to_arr(out_, s, "output");
  if (in.ndim != 1 || out.ndim != 1 || !all_contig({&in, &out}))
    throw py::value_error("Arguments must be C-contiguous vectors");
  if (in.dim(0) != out.dim(0))
    throw py::value_error("Input and output must have the same size");
  if (!in.same_type(out))
    throw py::value_error("Data types of arguments must be the same");
  if (!same_device({&in, &out})) throw py::value_error(kQueueMsg);
  if (!in.is_float())
    throw py::value_error(
        "Unsupported elemental data type. Expect single- or double- "
        "floating-point types.");
  use_device(in);
  wait_depends(s, depends);
  check_rc(kmd_iterative_merge_sort(in.dtype(), in.dim(0), in.data, out.data, s));
  return record(s, keep_of({&in, &out}));
}

py::tuple py_kmeans_lloyd_driver_ex(py::object X_, py::object w_,
                                    py::object init_, py::object lab_,
                                    py::object res_, double tol, bool verbose,
                                    size_t max_iter, const std::string &precision,
                                    py::object q, py::list depends) {
  kmd_stream_t s = resolve_stream(q);
  Arr X = to_arr(X_, s, "X_t"), w = to_arr(w_, s, "sample_weight");
  Arr init = to_arr(init_, s, "init_centroid_t");
  Arr lab = to_arr(lab_, s, "assignments_id");
  Arr res = to_arr(res_, s, "res_centroids_t");
  if (X.ndim != 2 || w.ndim != 1 || init.ndim != 2 || res.ndim != 2 ||
      lab.ndim != 1)
    throw py::value_error("Unsupported array dimensionalities");
  if (!all_contig({&X, &w, &init, &lab, &res}))
    throw py::value_error("All input arrays must be C-contiguous");
  if (!same_device({&X, &w, &init, &lab, &res}))
    throw py::value_error(kQueueMsg);
  const int64_t n_features = X.dim(0), n_samples = X.dim(1),
                n_clusters = init.dim(1);
  if (n_features != init.dim(0) || n_features != res.dim(0) ||
      n_clusters != res.dim(1) || n_samples != w.dim(0) ||
      n_samples != lab.dim(0))
    throw py::value_error("Array dimensions are not consistent");
  if (!X.same_type(w) || !X.same_type(init) || !X.same_type(res))
    throw py::value_error(
        "Sample coordinates, weights and centroids must have the same "
        "elemental data types");
  if (tol < 0.0) throw py::value_error("Tolerance must be non-negative");
  if (!X.is_float() || !lab.is_int())
    throw py::value_error("Unsupport elemental data type");
  int prec;
  if (precision == "fp32")
    prec = KMD_PREC_FP32;
  else if (precision == "bf16x3")
    prec = KMD_PREC_BF16X3;
  else if (precision == "bf16")
    prec = KMD_PREC_BF16;
  else if (precision == "fp8")
    prec = KMD_PREC_FP8;
  else
    throw py::value_error("precision must be 'fp32', 'bf16x3', 'bf16' or 'fp8'");
  use_device(X);
  wait_depends(s, depends);
  int64_t n_iters = 0, n_rescored = 0, launches = 0;
  double inertia = 0.0, kernel_ms = 0.0;
  int path = 0;
  int rc;
  {
    py::gil_scoped_release nogil;
    rc = kmd_kmeans_lloyd_driver_ex(
        X.dtype(), lab.itype(), prec, n_samples, n_features, n_clusters,
        X.data, w.data, init.data, lab.data, res.data, tol, verbose ? 1 : 0,
        (int64_t)max_iter, &n_iters, &inertia, &n_rescored, &kernel_ms,
        &launches, &path, verbose ? &print_to_python : nullptr, nullptr, s);
  }
  check_rc(rc);
  py::dict stats;
  stats["n_rescored"] = n_rescored;
  stats["fused_kernel_ms"] = kernel_ms;
  stats["fused_launches"] = launches;
  static const char *kPaths[] = {"generic",        "fp32-mfma",
                                 "bf16+rescore",   "bf16x3+rescore",
                                 "tiled-fp32",     "tiled-bf16+rescore",
                                 "tiled-bf16x3+rescore", "tiled-fp8+rescore"};
  stats["path"] = kPaths[(path >= 0 && path < 8) ? path : 0];
  return py::make_tuple((size_t)n_iters, inertia, stats);
}

// ---- estimator front-end: preparation and seeding -------------------------

std::pair<Event, Event> py_column_mean_var(py::object X_, py::object mean_,
                                           py::object var_, py::object q,
                                           py::list depends) {
  kmd_stream_t s = resolve_stream(q);
  Arr X = to_arr(X_, s, "X"), mean = to_arr(mean_, s, "mean");
  const bool has_var = !var_.is_none();
  Arr var = has_var ? to_arr(var_, s, "var") : Arr();
  if (X.ndim != 2 || mean.ndim != 1 || (has_var && var.ndim != 1))
    throw py::value_error("X must be a matrix, mean and var vectors");
  std::vector<const Arr *> all = {&X, &mean};
  if (has_var) all.push_back(&var);
  if (!all_contig(all)) throw py::value_error("All input arrays must be C-contiguous");
  if (!same_device(all)) throw py::value_error(kQueueMsg);
  if (!X.is_float() || mean.code != kDLFloat || mean.bits != 64 ||
      (has_var && (var.code != kDLFloat || var.bits != 64)))
    throw py::value_error("X must be float32/float64, mean and var float64");
  if (mean.dim(0) != X.dim(1) || (has_var && var.dim(0) != X.dim(1)))
    throw py::value_error("Array dimensions are not consistent");
  use_device(X);
  wait_depends(s, depends);
  check_rc(kmd_column_mean_var(X.dtype(), X.dim(0), X.dim(1), X.data,
                               static_cast<double *>(mean.data),
                               has_var ? static_cast<double *>(var.data) : nullptr,
                               s));
  return record(s, has_var ? keep_of({&X, &mean, &var}) : keep_of({&X, &mean}));
}

std::pair<Event, Event> py_transpose(py::object src_, py::object dst_,
                                     py::object shift_, py::object q,
                                     py::list depends) {
  kmd_stream_t s = resolve_stream(q);
  Arr src = to_arr(src_, s, "src"), dst = to_arr(dst_, s, "dst");
  const bool has_shift = !shift_.is_none();
  Arr shift = has_shift ? to_arr(shift_, s, "shift") : Arr();
  if (src.ndim != 2 || dst.ndim != 2 || (has_shift && shift.ndim != 1))
    throw py::value_error("src and dst must be matrices, shift a vector");
  std::vector<const Arr *> all = {&src, &dst};
  if (has_shift) all.push_back(&shift);
  if (!all_contig(all)) throw py::value_error("All input arrays must be C-contiguous");
  if (!same_device(all)) throw py::value_error(kQueueMsg);
  if (!src.is_float() || !src.same_type(dst) ||
      (has_shift && (shift.code != kDLFloat || shift.bits != 64)))
    throw py::value_error("src and dst must share a float dtype, shift float64");
  if (dst.dim(0) != src.dim(1) || dst.dim(1) != src.dim(0) ||
      (has_shift && shift.dim(0) != src.dim(1)))
    throw py::value_error("Array dimensions are not consistent");
  use_device(src);
  wait_depends(s, depends);
  check_rc(kmd_transpose(src.dtype(), src.dim(0), src.dim(1), src.data,
                         has_shift ? static_cast<const double *>(shift.data) : nullptr,
                         dst.data, s));
  return record(s, has_shift ? keep_of({&src, &dst, &shift}) : keep_of({&src, &dst}));
}

std::pair<Event, Event> py_gather_columns(py::object X_, py::object idx_,
                                          py::object out_, py::object q,
                                          py::list depends) {
  kmd_stream_t s = resolve_stream(q);
  Arr X = to_arr(X_, s, "X_t"), idx = to_arr(idx_, s, "indices"),
      out = to_arr(out_, s, "out_t");
  if (X.ndim != 2 || idx.ndim != 1 || out.ndim != 2)
    throw py::value_error("X_t and out_t must be matrices, indices a vector");
  if (!all_contig({&X, &idx, &out}))
    throw py::value_error("All input arrays must be C-contiguous");
  if (!same_device({&X, &idx, &out})) throw py::value_error(kQueueMsg);
  if (!X.is_float() || !X.same_type(out) || idx.code != kDLInt || idx.bits != 64)
    throw py::value_error("X_t and out_t must share a float dtype, indices int64");
  if (out.dim(0) != X.dim(0) || out.dim(1) != idx.dim(0))
    throw py::value_error("Array dimensions are not consistent");
  use_device(X);
  wait_depends(s, depends);
  check_rc(kmd_gather_columns(X.dtype(), X.dim(1), X.dim(0), idx.dim(0), X.data,
                              static_cast<const int64_t *>(idx.data), out.data, s));
  return record(s, keep_of({&X, &idx, &out}));
}

std::pair<Event, Event> py_kmeans_plusplus(py::object X_, py::object w_,
                                           py::array_t<double, py::array::c_style |
                                                                   py::array::forcecast>
                                               uniforms,
                                           py::object C_, py::object idx_,
                                           py::object q, py::list depends) {
  kmd_stream_t s = resolve_stream(q);
  Arr X = to_arr(X_, s, "X_t"), w = to_arr(w_, s, "sample_weight"),
      C = to_arr(C_, s, "centers_t"), idx = to_arr(idx_, s, "indices");
  if (X.ndim != 2 || w.ndim != 1 || C.ndim != 2 || idx.ndim != 1)
    throw py::value_error("Array dimensionality is not consistent");
  if (!all_contig({&X, &w, &C, &idx}))
    throw py::value_error("All input arrays must be C-contiguous");
  if (!same_device({&X, &w, &C, &idx})) throw py::value_error(kQueueMsg);
  if (!X.is_float() || !X.same_type(w) || !X.same_type(C) ||
      idx.code != kDLInt || idx.bits != 64)
    throw py::value_error("X_t, sample_weight, centers_t must share a float "
                          "dtype; indices must be int64");
  const int64_t d = X.dim(0), n = X.dim(1), k = C.dim(1);
  if (C.dim(0) != d || w.dim(0) != n || idx.dim(0) != k)
    throw py::value_error("Array dimensions are not consistent");
  if (uniforms.ndim() != 2 || uniforms.shape(0) < 1)
    throw py::value_error("uniforms must be a (1 + (k - 1), T) array: row 0 "
                          "holds the first center's draw");
  const int64_t T = uniforms.shape(1);
  if (uniforms.shape(0) != k)
    throw py::value_error("uniforms must have n_clusters rows");
  // flatten to [u0, row 1, row 2, ...]
  std::vector<double> u((size_t)(1 + (k - 1) * T));
  const double *up = uniforms.data();
  u[0] = up[0];
  for (int64_t c = 1; c < k; ++c)
    for (int64_t j = 0; j < T; ++j) u[(size_t)(1 + (c - 1) * T + j)] = up[c * T + j];
  use_device(X);
  wait_depends(s, depends);
  {
    py::gil_scoped_release nogil;
    check_rc(kmd_kmeans_plusplus(X.dtype(), n, d, k, T, X.data, w.data, u.data(),
                                 C.data, static_cast<int64_t *>(idx.data), s));
  }
  return record(s, keep_of({&X, &w, &C, &idx}));
}

}  // namespace

py::tuple py_kmeans_lloyd_batched(py::list fits, const std::string &precision,
                                  py::object q, int graph_replays,
                                  py::list depends) {
  if (precision != "fp32" && precision != "bf16")
    throw py::value_error("batched fits support precision 'fp32' or 'bf16'");
  const int prec = precision == "bf16" ? KMD_PREC_BF16 : KMD_PREC_FP32;
  kmd_stream_t s = resolve_stream(q);
  std::vector<kmd_fit> v;
  std::vector<Arr> keep;
  keep.reserve(5 * fits.size());
  for (py::handle h : fits) {
    py::tuple t = py::reinterpret_borrow<py::tuple>(h);
    if (t.size() != 7)
      throw py::value_error(
          "each fit is (X_t, sample_weight, init_centroid_t, assignments_id, "
          "res_centroids_t, tol, max_iter)");
    Arr X = to_arr(t[0], s, "X_t"), w = to_arr(t[1], s, "sample_weight");
    Arr init = to_arr(t[2], s, "init_centroid_t");
    Arr lab = to_arr(t[3], s, "assignments_id");
    Arr res = to_arr(t[4], s, "res_centroids_t");
    if (X.ndim != 2 || w.ndim != 1 || init.ndim != 2 || res.ndim != 2 ||
        lab.ndim != 1)
      throw py::value_error("Unsupported array dimensionalities");
    if (!all_contig({&X, &w, &init, &lab, &res}))
      throw py::value_error("All input arrays must be C-contiguous");
    if (!same_device({&X, &w, &init, &lab, &res}))
      throw py::value_error(kQueueMsg);
    const int64_t d = X.dim(0), n = X.dim(1), k = init.dim(1);
    if (d != init.dim(0) || d != res.dim(0) || k != res.dim(1) ||
        n != w.dim(0) || n != lab.dim(0))
      throw py::value_error("Array dimensions are not consistent");
    if (!X.is_float() || X.dtype() != KMD_F32 || !X.same_type(w) ||
        !X.same_type(init) || !X.same_type(res) || !lab.is_int() ||
        lab.itype() != KMD_I32)
      throw py::value_error("batched fits take float32 data and int32 labels");
    kmd_fit f{};
    f.n_samples = n;
    f.n_features = d;
    f.n_clusters = k;
    f.X_t = X.data;
    f.sample_weight = w.data;
    f.init_centroids_t = init.data;
    f.assignments_id = lab.data;
    f.res_centroids_t = res.data;
    f.tol = t[5].cast<double>();
    f.max_iter = t[6].cast<int64_t>();
    if (f.tol < 0.0) throw py::value_error("Tolerance must be non-negative");
    if (f.max_iter < 0) throw py::value_error("max_iter must be non-negative");
    v.push_back(f);
    use_device(X);
    keep.push_back(std::move(X));
    keep.push_back(std::move(w));
    keep.push_back(std::move(init));
    keep.push_back(std::move(lab));
    keep.push_back(std::move(res));
  }
  wait_depends(s, depends);
  int64_t launches = 0;
  int shape = 0;
  int rc;
  {
    py::gil_scoped_release nogil;
    rc = kmd_kmeans_lloyd_batched((int64_t)v.size(), v.data(), prec,
                                  graph_replays, &launches, &shape, s);
  }
  check_rc(rc);
  py::list out;
  for (const kmd_fit &f : v) out.append(py::make_tuple((size_t)f.n_iters, f.inertia));
  py::dict stats;
  stats["graph_launches"] = launches;
  stats["assign_shape"] = shape;
  return py::make_tuple(out, stats);
}

// Continuous-batching engine: fits join a running set and advance together.
class Engine {
 public:
  explicit Engine(py::object q, const std::string &precision) : q_(q) {
    if (precision != "fp32" && precision != "bf16")
      throw py::value_error("engine precision must be 'fp32' or 'bf16'");
    s_ = resolve_stream(q);
    check_rc(kmd_engine_create(&e_, s_));
    const int rc = kmd_engine_set_precision(
        e_, precision == "bf16" ? KMD_PREC_BF16 : KMD_PREC_FP32);
    if (rc != KMD_OK) {
      kmd_engine_destroy(e_);
      e_ = nullptr;
      check_rc(rc);
    }
  }
  ~Engine() {
    if (e_) kmd_engine_destroy(e_);
  }
  int64_t add(py::object X_, py::object w_, py::object init_, py::object lab_,
              py::object res_, double tol, int64_t max_iter) {
    auto ent = std::make_unique<Entry>();
    ent->keep.push_back(to_arr(X_, s_, "X_t"));
    ent->keep.push_back(to_arr(w_, s_, "sample_weight"));
    ent->keep.push_back(to_arr(init_, s_, "init_centroid_t"));
    ent->keep.push_back(to_arr(lab_, s_, "assignments_id"));
    ent->keep.push_back(to_arr(res_, s_, "res_centroids_t"));
    Arr &X = ent->keep[0], &w = ent->keep[1], &init = ent->keep[2],
        &lab = ent->keep[3], &res = ent->keep[4];
    if (X.ndim != 2 || w.ndim != 1 || init.ndim != 2 || res.ndim != 2 ||
        lab.ndim != 1)
      throw py::value_error("Unsupported array dimensionalities");
    if (!all_contig({&X, &w, &init, &lab, &res}))
      throw py::value_error("All input arrays must be C-contiguous");
    if (!same_device({&X, &w, &init, &lab, &res}))
      throw py::value_error(kQueueMsg);
    const int64_t d = X.dim(0), n = X.dim(1), k = init.dim(1);
    if (d != init.dim(0) || d != res.dim(0) || k != res.dim(1) ||
        n != w.dim(0) || n != lab.dim(0))
      throw py::value_error("Array dimensions are not consistent");
    if (!X.is_float() || X.dtype() != KMD_F32 || !X.same_type(w) ||
        !X.same_type(init) || !X.same_type(res) || !lab.is_int() ||
        lab.itype() != KMD_I32)
      throw py::value_error("engine fits take float32 data and int32 labels");
    if (tol < 0.0) throw py::value_error("Tolerance must be non-negative");
    if (max_iter < 0) throw py::value_error("max_iter must be non-negative");
    kmd_fit &f = ent->fit;
    f.n_samples = n;
    f.n_features = d;
    f.n_clusters = k;
    f.X_t = X.data;
    f.sample_weight = w.data;
    f.init_centroids_t = init.data;
    f.assignments_id = lab.data;
    f.res_centroids_t = res.data;
    f.tol = tol;
    f.max_iter = max_iter;
    use_device(X);
    int64_t id = -1;
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = kmd_engine_add(e_, &ent->fit, &id);
    }
    check_rc(rc);
    entries_[id] = std::move(ent);
    return id;
  }
  int64_t step(int iterations, bool use_graph) {
    int64_t active = 0;
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = kmd_engine_step(e_, iterations, use_graph ? 1 : 0, &active);
    }
    check_rc(rc);
    return active;
  }
  py::list take_finished() {
    py::list out;
    for (;;) {
      int64_t ids[64], cnt = 0;
      check_rc(kmd_engine_take_finished(e_, ids, 64, &cnt));
      for (int64_t i = 0; i < cnt; ++i) {
        auto it = entries_.find(ids[i]);
        if (it == entries_.end()) continue;
        out.append(py::make_tuple(ids[i], (size_t)it->second->fit.n_iters,
                                  it->second->fit.inertia));
        entries_.erase(it);
      }
      if (cnt < 64) break;
    }
    return out;
  }
  int64_t resident_bytes() {
    int64_t b = 0;
    check_rc(kmd_engine_resident_bytes(e_, &b));
    return b;
  }
  static int64_t footprint(int64_t n, int64_t d, int64_t k) {
    int64_t b = 0;
    check_rc(kmd_engine_footprint(n, d, k, &b));
    return b;
  }
  size_t size() const { return entries_.size(); }

 private:
  struct Entry {
    kmd_fit fit{};
    std::vector<Arr> keep;
  };
  py::object q_;
  kmd_stream_t s_ = nullptr;
  kmd_engine_t e_ = nullptr;
  std::map<int64_t, std::unique_ptr<Entry>> entries_;
};

py::dict py_autotune(py::object X_, py::object w_, py::object init_,
                     const std::string &precision, int iters, py::object q) {
  kmd_stream_t s = resolve_stream(q);
  Arr X = to_arr(X_, s, "X_t"), w = to_arr(w_, s, "sample_weight");
  Arr init = to_arr(init_, s, "init_centroid_t");
  if (X.ndim != 2 || w.ndim != 1 || init.ndim != 2)
    throw py::value_error("Unsupported array dimensionalities");
  if (!all_contig({&X, &w, &init}))
    throw py::value_error("All input arrays must be C-contiguous");
  if (!X.is_float() || !X.same_type(w) || !X.same_type(init))
    throw py::value_error("Unsupport elemental data type");
  const int64_t d = X.dim(0), n = X.dim(1), k = init.dim(1);
  if (d != init.dim(0) || n != w.dim(0))
    throw py::value_error("Array dimensions are not consistent");
  int prec = precision == "fp32" ? KMD_PREC_FP32
             : precision == "bf16x3" ? KMD_PREC_BF16X3
             : precision == "bf16" ? KMD_PREC_BF16
             : precision == "fp8" ? KMD_PREC_FP8 : -1;
  if (prec < 0)
    throw py::value_error("precision must be 'fp32', 'bf16x3', 'bf16' or 'fp8'");
  use_device(X);
  constexpr int kCap = 64;
  int bp = 0, bb = 0, bs = 0;
  double bms = 0.0;
  int cp[kCap], cb[kCap], cs[kCap];
  double cm[kCap];
  int64_t nc = 0;
  int rc;
  {
    py::gil_scoped_release nogil;
    rc = kmd_autotune(X.dtype(), KMD_I32, prec, n, d, k, X.data, w.data, init.data,
                      iters, &bp, &bb, &bs, &bms, cp, cb, cs, cm, kCap, &nc, s);
  }
  check_rc(rc);
  static const char *kPath[] = {"generic", "lds-resident", "tiled"};
  py::list cands;
  for (int64_t i = 0; i < std::min<int64_t>(nc, kCap); ++i)
    cands.append(py::make_tuple(kPath[cp[i]], cb[i], cs[i], cm[i]));
  py::dict out;
  out["path"] = kPath[bp];
  out["blocks_per_cu"] = bb;
  out["shape"] = bs;
  out["ms"] = bms;
  out["candidates"] = cands;
  return out;
}

PYBIND11_MODULE(_kmeans_dpcpp_amd, m) {
  m.doc() =
      "MI355X (gfx950) HIP implementation of the Lloyd k-means entry points "
      "of kmeans_dpcpp.";

  py::class_<Event>(m, "Event")
      .def(py::init<>())
      .def("wait", &Event::wait, "Block until the work is complete and release "
                                 "the kept-alive arguments.")
      .def_property_readonly("is_complete", &Event::is_complete)
      .def("elapsed_ms", &Event::elapsed_ms, py::arg("stop"));

  m.def("version", []() { return std::string(kmd_version()); });
  m.def("device_count", []() {
    int c = 0;
    check_rc(kmd_device_count(&c));
    return c;
  });
  m.def("_stream_create", []() {
    kmd_stream_t s;
    check_rc(kmd_stream_create(&s));
    return reinterpret_cast<intptr_t>(s);
  });
  m.def("_stream_destroy", [](intptr_t s) {
    check_rc(kmd_stream_destroy(reinterpret_cast<kmd_stream_t>(s)));
  });
  m.def("_stream_synchronize", [](py::object q) {
    kmd_stream_t s = resolve_stream(q);
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = kmd_stream_synchronize(s);
    }
    check_rc(rc);
  });
  m.def("record_event", [](py::object q) { return record(resolve_stream(q), py::list()).second; },
        py::arg("sycl_queue"));

  m.def("broadcast_divide", &py_broadcast_divide,
        "broadcast_divide(divident, divisor, work_group_size, sycl_queue, "
        "depends=[]) evaluates `divident /= divisor` for a matrix and a vector",
        py::arg("divident"), py::arg("divisor"), py::arg("work_group_size"),
        py::arg("sycl_queue"), py::arg("depends") = py::list());
  m.def("half_l2_norm_squared", &py_half_l2_norm_squared,
        "half_l2_norm_squared(centroids, centroids_half_l2_norm_squared, "
        "work_group_size, sycl_queue, depends=[]) computes column-wise half of "
        "the squared L2 norm",
        py::arg("centroids"), py::arg("centroids_half_l2_norm_squared"),
        py::arg("work_group_size"), py::arg("sycl_queue"),
        py::arg("depends") = py::list());
  m.def("reduce_centroids_data", &py_reduce_centroids_data,
        "Sums private copies of centroid sums and cluster sizes and lists "
        "empty clusters",
        py::arg("cluster_sizes_private_copies"),
        py::arg("centroids_t_private_copies"), py::arg("out_cluster_sizes"),
        py::arg("out_centroids_t"), py::arg("out_empty_clusters_list"),
        py::arg("out_n_empty_clusters"), py::arg("work_group_size"),
        py::arg("sycl_queue"), py::arg("depends") = py::list());
  m.def("compute_threshold", &py_compute_threshold,
        "compute_threshold(data, topk, threshold, sycl_queue, depends=[]) "
        "puts the topk-th largest element of data in threshold",
        py::arg("data"), py::arg("topk"), py::arg("threshold"),
        py::arg("sycl_queue"), py::arg("depends") = py::list());
  m.def("select_samples_far_from_centroid",
        &py_select_samples_far_from_centroid,
        "Writes ids of samples farther than threshold at the front of "
        "selected_samples_idx and ids at exactly threshold at its back",
        py::arg("n_selected"), py::arg("distance_to_centroid"),
        py::arg("threshold"), py::arg("selected_samples_idx"),
        py::arg("n_selected_gt_threshold"), py::arg("n_selected_eq_threshold"),
        py::arg("work_group_size"), py::arg("sycl_queue"),
        py::arg("depends") = py::list());
  m.def("relocate_empty_clusters", &py_relocate_empty_clusters,
        "Relocates empty clusters; modifies centroid_t, cluster_sizes, "
        "per_sample_inertia",
        py::arg("n_empty_clusters"), py::arg("X_t"), py::arg("sample_weights"),
        py::arg("assignment_id"), py::arg("empty_clusters_list"),
        py::arg("sq_dist_to_nearest_centroid"), py::arg("centroid_t"),
        py::arg("cluster_sizes"), py::arg("per_sample_inertia"),
        py::arg("work_group_size"), py::arg("sycl_queue"),
        py::arg("depends") = py::list());
  m.def("compute_centroid_shifts_squared", &py_compute_centroid_shifts_squared,
        "Computes np.sum(np.square(old_centroid_t - new_centroid_t), axis=0)",
        py::arg("centroid_t"), py::arg("new_centroid_t"),
        py::arg("out_centroid_shifts"), py::arg("sycl_queue"),
        py::arg("depends") = py::list());
  m.def("compute_centroid_to_sample_distances", &py_compute_distances,
        "Distances from centroids (n_features, n_clusters) to samples "
        "(n_features, n_samples) into (n_clusters, n_samples)",
        py::arg("X_t"), py::arg("centroid_t"), py::arg("euclidean_distances_t"),
        py::arg("work_group_size"), py::arg("centroids_window_height"),
        py::arg("sycl_queue"), py::arg("depends") = py::list());
  m.def("assignment", &py_assignment,
        "Compute assignment of samples to nearest centroids.", py::arg("X_t"),
        py::arg("centroids_t"), py::arg("centroids_half_l2_norm"),
        py::arg("assignment_id"), py::arg("centroids_window_height"),
        py::arg("work_group_size"), py::arg("sycl_queue"),
        py::arg("depends") = py::list());
  m.def("compute_inertia", &py_compute_inertia,
        "Computes per sample inertia given assignment IDs", py::arg("X_t"),
        py::arg("sample_weight"), py::arg("centroid_t"),
        py::arg("assignment_id"), py::arg("out_per_sample_inertia"),
        py::arg("work_group_size"), py::arg("sycl_queue"),
        py::arg("depends") = py::list());
  m.def("reduce_vector_blocking", &py_reduce_vector_blocking,
        "Synchronously compute the total value of elements in the vector",
        py::arg("data"), py::arg("sycl_queue"), py::arg("depends") = py::list());
  m.def("fused_lloyd_single_step", &py_fused_lloyd_single_step,
        "Perform single step of Lloyd' algorithm for KMeans problem",
        py::arg("X_t"), py::arg("sample_weight"), py::arg("centroids_t"),
        py::arg("centroids_half_l2_norm"), py::arg("assignments_idx"),
        py::arg("new_centroids_t_private_copies"),
        py::arg("cluster_sizes_private_copies"),
        py::arg("centroids_window_height"), py::arg("work_group_size"),
        py::arg("sycl_queue"), py::arg("depends") = py::list());
  m.def("compute_number_of_private_copies",
        &py_compute_number_of_private_copies,
        "Number of private copies for a max-cache-occupancy fraction.",
        py::arg("array"), py::arg("n_samples"), py::arg("n_features"),
        py::arg("n_clusters"),
        py::arg("centroids_private_copies_max_cache_occupancy"),
        py::arg("work_group_size"));
  m.def("kmeans_lloyd_driver", &py_kmeans_lloyd_driver,
        "Lloyd's refinement. Returns (n_iters, array of shape (1,) with the "
        "total inertia). Array init_centroid_t is overwritten.",
        py::arg("X_t"), py::arg("sample_weight"), py::arg("init_centroid_t"),
        py::arg("assignments_id"), py::arg("res_centroids_t"), py::arg("tol"),
        py::arg("verbose"), py::arg("max_iter"),
        py::arg("centroids_window_height"), py::arg("work_group_size"),
        py::arg("centroids_private_copies_max_cache_occupancy"),
        py::arg("sycl_queue"), py::arg("depends") = py::list());
  m.def("iterative_merge_sort", &py_iterative_merge_sort,
        "Stable ascending sort of a float vector (NaN last) into output",
        py::arg("data"), py::arg("output"), py::arg("sycl_queue"),
        py::arg("depends") = py::list());
  py::class_<Engine>(m, "Engine",
                     "Continuous batching of Lloyd fits on one GPU: add() joins a "
                     "fit to the running set, step() advances every active fit, "
                     "take_finished() returns [(id, n_iters, inertia)] of fits that "
                     "stopped.  Each fit's results equal kmeans_lloyd_driver's for it "
                     "alone.  precision 'bf16': the assignment takes the cross term in "
                     "BF16 with a certified FP32 re-score (the same results).")
      .def(py::init<py::object, const std::string &>(), py::arg("sycl_queue") = py::none(),
           py::arg("precision") = "fp32")
      .def("add", &Engine::add, py::arg("X_t"), py::arg("sample_weight"),
           py::arg("init_centroid_t"), py::arg("assignments_id"),
           py::arg("res_centroids_t"), py::arg("tol"), py::arg("max_iter"))
      .def("step", &Engine::step, py::arg("iterations") = 1,
           py::arg("use_graph") = false)
      .def("take_finished", &Engine::take_finished)
      .def("resident_bytes", &Engine::resident_bytes)
      .def("__len__", &Engine::size)
      .def_static("footprint", &Engine::footprint, py::arg("n_samples"),
                  py::arg("n_features"), py::arg("n_clusters"));
  m.def("autotune", &py_autotune,
        "Time the driver's launch plans (LDS-resident or tiled centroids x "
        "launch shape of the LDS-resident step x persistent blocks per CU) for "
        "this problem and keep the fastest for later fits of its shape; results "
        "never depend on the plan.  Returns {'path', 'blocks_per_cu', 'shape', "
        "'ms', 'candidates': [(path, blocks_per_cu, shape, ms)]}.",
        py::arg("X_t"), py::arg("sample_weight"), py::arg("init_centroid_t"),
        py::arg("precision") = "fp32", py::arg("iters") = 3,
        py::arg("sycl_queue") = py::none());
  m.def("fp8_roundtrip",
        [](py::object in_, py::object out_, py::object q) {
          kmd_stream_t s = resolve_stream(q);
          Arr in = to_arr(in_, s, "in"), out = to_arr(out_, s, "out");
          if (in.ndim != 1 || out.ndim != 1 || in.dim(0) != out.dim(0) ||
              !in.is_float() || in.dtype() != KMD_F32 || !in.same_type(out))
            throw py::value_error("in and out must be float32 vectors of one length");
          use_device(in);
          check_rc(kmd_fp8_roundtrip(in.dim(0), static_cast<const float *>(in.data),
                                     static_cast<float *>(out.data), s));
        },
        "out = in rounded to FP8 e4m3 and back with the FP8 kernels' conversion.",
        py::arg("in"), py::arg("out"), py::arg("sycl_queue") = py::none());
  m.def("set_tuning",
        [](int path, int blocks_per_cu, int shape) {
          check_rc(kmd_set_tuning(path, blocks_per_cu, shape));
        },
        "Override the launch plan of driver calls: path 0 automatic, 1 "
        "LDS-resident, 2 tiled (-1 clears the override); blocks_per_cu 0 = "
        "occupancy limit; shape 0 (16 waves x 1 column block per wave), 1 (12 x "
        "2) or 2 (8 x 2) for the LDS-resident step.",
        py::arg("path"), py::arg("blocks_per_cu") = 0, py::arg("shape") = 0);
  m.def("kmeans_lloyd_batched", &py_kmeans_lloyd_batched,
        "Many independent Lloyd fits at once: fits is a list of (X_t, "
        "sample_weight, init_centroid_t, assignments_id, res_centroids_t, tol, "
        "max_iter).  Each fit's results equal kmeans_lloyd_driver's for it "
        "alone.  Returns ([(n_iters, inertia), ...], stats).",
        py::arg("fits"), py::arg("precision") = "fp32",
        py::arg("sycl_queue") = py::none(), py::arg("graph_replays") = 4,
        py::arg("depends") = py::list());
  m.def("kmeans_lloyd_driver_ex", &py_kmeans_lloyd_driver_ex,
        "Lloyd fit with a selectable cross-term precision ('fp32', 'bf16x3', "
        "'bf16'). Returns (n_iters, total_inertia, stats) where stats holds "
        "n_rescored, fused_kernel_ms (summed HIP-event time of the fused-step "
        "kernel) and fused_launches.",
        py::arg("X_t"), py::arg("sample_weight"), py::arg("init_centroid_t"),
        py::arg("assignments_id"), py::arg("res_centroids_t"), py::arg("tol"),
        py::arg("verbose"), py::arg("max_iter"), py::arg("precision"),
        py::arg("sycl_queue"), py::arg("depends") = py::list());
  m.def("column_mean_var", &py_column_mean_var,
        "column_mean_var(X, mean, var=None, sycl_queue=None, depends=[]): "
        "per-feature mean and variance (float64 outputs) of a row-major "
        "(n_samples, n_features) X.",
        py::arg("X"), py::arg("mean"), py::arg("var") = py::none(),
        py::arg("sycl_queue") = py::none(), py::arg("depends") = py::list());
  m.def("transpose", &py_transpose,
        "transpose(src, dst, shift=None, sycl_queue=None, depends=[]): "
        "dst = src.T - shift (shift: float64 per column of src, or None).",
        py::arg("src"), py::arg("dst"), py::arg("shift") = py::none(),
        py::arg("sycl_queue") = py::none(), py::arg("depends") = py::list());
  m.def("gather_columns", &py_gather_columns,
        "gather_columns(X_t, indices, out_t, sycl_queue=None, depends=[]): "
        "out_t[:, j] = X_t[:, indices[j]] (indices int64 on the device).",
        py::arg("X_t"), py::arg("indices"), py::arg("out_t"),
        py::arg("sycl_queue") = py::none(), py::arg("depends") = py::list());
  m.def("kmeans_plusplus_t", &py_kmeans_plusplus,
        "kmeans_plusplus_t(X_t, sample_weight, uniforms, centers_t, indices, "
        "sycl_queue=None, depends=[]): greedy k-means++ (scikit-learn's "
        "algorithm) with the random draws given: uniforms is a host float64 "
        "array (n_clusters, n_local_trials) whose [0, 0] picks the first "
        "center and row c the candidates of center c.",
        py::arg("X_t"), py::arg("sample_weight"), py::arg("uniforms"),
        py::arg("centers_t"), py::arg("indices"),
        py::arg("sycl_queue") = py::none(), py::arg("depends") = py::list());
}
