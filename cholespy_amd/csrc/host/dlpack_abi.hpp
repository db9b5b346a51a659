// Minimal declaration of the DLPack exchange ABI (legacy, unversioned
// "dltensor" capsules), the protocol NumPy, PyTorch (CPU and ROCm) and other
// frameworks export through __dlpack__(). The struct layouts are fixed by the
// DLPack standard (dmlc/dlpack, dlpack.h v0.8); only what the binding reads
// is declared.
#pragma once
#include <cstdint>

namespace cholespy_amd {
namespace dl {

enum DeviceType : int32_t {
    kCPU = 1,
    kCUDA = 2,
    kCUDAHost = 3,
    kROCM = 10,
    kROCMHost = 11,
};

enum TypeCode : uint8_t { kInt = 0, kUInt = 1, kFloat = 2, kBfloat = 4 };

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
    int64_t *strides;   // in elements; NULL means compact row-major
    uint64_t byte_offset;
};

struct ManagedTensor {
    Tensor dl_tensor;
    void *manager_ctx;
    void (*deleter)(ManagedTensor *self);
};

}  // namespace dl
}  // namespace cholespy_amd

namespace cholespy_amd {
namespace dl {

// DLPack >= 1.0 capsules ("dltensor_versioned"), requested with
// __dlpack__(max_version=(1, 0)); NumPy only exports read-only arrays this way.
struct PackVersion {
    uint32_t major;
    uint32_t minor;
};

struct ManagedTensorVersioned {
    PackVersion version;
    void *manager_ctx;
    void (*deleter)(ManagedTensorVersioned *self);
    uint64_t flags;
    Tensor dl_tensor;
};

}  // namespace dl
}  // namespace cholespy_amd
