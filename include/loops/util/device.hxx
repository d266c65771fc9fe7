/**
 * @file device.hxx
 * @brief Device selection and properties (reference include/loops/util/
 * device.hxx:23-52).  multi_processor_count() is the CU count (256 on
 * MI355X).
 */
#pragma once

#include <hip/hip_runtime_api.h>

namespace loops {
namespace device {

typedef int device_id_t;

inline void set(device_id_t ordinal) { (void)hipSetDevice(ordinal); }

inline device_id_t get() {
  device_id_t ordinal = 0;
  (void)hipGetDevice(&ordinal);
  return ordinal;
}

struct properties_t {
  typedef hipDeviceProp_t device_properties_t;
  device_properties_t properties;
  device_id_t ordinal;
  properties_t() : ordinal(device::get()) {
    (void)hipGetDeviceProperties(&properties, ordinal);
  }
  int multi_processor_count() { return properties.multiProcessorCount; }
};

}  // namespace device
}  // namespace loops
