/**
 * @file math.hxx
 * @brief Integer helpers (reference include/loops/util/math.hxx:26-31).
 */
#pragma once

#include <hip/hip_runtime.h>

namespace loops {
namespace math {

/// ceil(n / d) for non-negative integers.
template <typename numerator_t, typename denominator_t>
__host__ __device__ constexpr numerator_t ceil_div(numerator_t const& n,
                                                   denominator_t const& d) {
  return static_cast<numerator_t>(n / d + (n % d != 0 ? 1 : 0));
}

}  // namespace math
}  // namespace loops
