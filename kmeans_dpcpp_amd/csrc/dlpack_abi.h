// The DLPack exchange structures (DLPack 0.8 ABI, data-apis.org), declared
// here because the boundary only reads them.  Field order and types must stay
// exactly as the standard fixes them.
#pragma once
#include <stdint.h>

namespace kmd_dlpack {

enum DeviceType : int32_t {
  kDLCPU = 1,
  kDLCUDA = 2,
  kDLCUDAHost = 3,
  kDLROCM = 10,
  kDLROCMHost = 11,
};

enum TypeCode : uint8_t { kDLInt = 0, kDLUInt = 1, kDLFloat = 2, kDLBfloat = 4 };

struct Device {
  int32_t device_type;
  int32_t device_id;
};

struct DataType {
  uint8_t code;
  uint8_t bits;
  uint16_t lanes;
};

struct Tensor {
  void *data;
  Device device;
  int32_t ndim;
  DataType dtype;
  int64_t *shape;
  int64_t *strides;  // in elements; NULL means compact row-major
  uint64_t byte_offset;
};

struct ManagedTensor {
  Tensor dl_tensor;
  void *manager_ctx;
  void (*deleter)(ManagedTensor *self);
};

}  // namespace kmd_dlpack
