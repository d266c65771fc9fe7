/**
 * @file stride_ranges.hxx
 * @brief Grid-, block- and custom-stride ranges (reference
 * include/loops/stride_ranges.hxx:17-63).  On CDNA a "grid stride" is
 * gridDim.x * blockDim.x work-items, i.e. (#workgroups x 64-wide waves).
 */
#pragma once

#include <loops/range.hxx>

namespace loops {

template <typename T>
using step_range_t = typename range_proxy<T>::step_range_proxy;

template <typename T>
__device__ __forceinline__ step_range_t<T> grid_stride_range(T begin, T end) {
  begin += T(blockDim.x) * T(blockIdx.x) + T(threadIdx.x);
  return range(begin, end).step(T(gridDim.x) * T(blockDim.x));
}

template <typename T>
__device__ __forceinline__ step_range_t<T> block_stride_range(T begin, T end) {
  return range(begin, end).step(T(blockDim.x));
}

template <typename T>
__host__ __device__ __forceinline__ step_range_t<T> custom_stride_range(
    T begin, T end, T stride) {
  return range(begin, end).step(stride);
}

}  // namespace loops
