/**
 * @file equal.hxx
 * @brief Device-vs-host comparison (reference include/loops/util/equal.hxx:
 * 40-62): copies n device values to the host and counts the positions where
 * error_op(device, host) is true, optionally printing them.
 */
#pragma once

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstddef>
#include <iomanip>
#include <iostream>
#include <limits>
#include <vector>

namespace loops {
namespace util {

template <typename type_t, typename comp_t>
std::size_t equal(const type_t* d_ptr, const type_t* h_ptr, const std::size_t n,
                  comp_t error_op, const bool verbose = false) {
  std::vector<type_t> d(n);
  if (n) (void)hipMemcpy(d.data(), d_ptr, n * sizeof(type_t), hipMemcpyDeviceToHost);
  std::size_t errors = 0;
  for (std::size_t i = 0; i < n; ++i) {
    if (error_op(d[i], h_ptr[i])) {
      if (verbose)
        std::cout << "Error[" << i << "]: " << std::setw(10) << std::fixed
                  << std::setprecision(std::numeric_limits<type_t>::digits10)
                  << d[i] << " != " << std::setw(10) << h_ptr[i] << std::endl;
      ++errors;
    }
  }
  return errors;
}

template <typename type_t>
std::size_t equal(const type_t* d_ptr, const type_t* h_ptr, const std::size_t n) {
  return equal(d_ptr, h_ptr, n, [](type_t a, type_t b) { return a != b; });
}

}  // namespace util
}  // namespace loops
