/**
 * @file memory.hxx
 * @brief Memory spaces and raw-pointer access (reference
 * include/loops/memory.hxx:33-65).  device = hipMalloc'd HBM, host =
 * pageable host memory, managed = hipMallocManaged.
 */
#pragma once

#include <hip/hip_runtime.h>

namespace loops {
namespace memory {

enum memory_space_t { device, host, managed };

/// Device pointer wrapper whose .get() returns the raw pointer, so code
/// written as `v.data().get()` keeps working.
template <typename type_t>
struct device_ptr {
  type_t* ptr;
  __host__ __device__ type_t* get() const { return ptr; }
  __host__ __device__ operator type_t*() const { return ptr; }
};

template <typename type_t>
__host__ __device__ inline type_t* raw_pointer_cast(device_ptr<type_t> p) {
  return p.get();
}

template <typename type_t>
__host__ __device__ inline type_t* raw_pointer_cast(type_t* p) {
  return p;
}

}  // namespace memory

using memory::memory_space_t;

}  // namespace loops
