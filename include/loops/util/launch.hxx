/**
 * @file launch.hxx
 * @brief Kernel launch helpers (reference include/loops/util/launch.hxx:
 * 31-74): non_cooperative = a <<<grid, block, 0, stream>>> launch;
 * cooperative = hipLaunchCooperativeKernel with the argument addresses.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

namespace loops {
namespace launch {
namespace detail {
inline void for_each_argument_address(void**) {}
template <typename arg_t, typename... args_t>
inline void for_each_argument_address(void** out, arg_t&& arg, args_t&&... args) {
  out[0] = const_cast<void*>(static_cast<const void*>(&arg));
  for_each_argument_address(out + 1, std::forward<args_t>(args)...);
}
}  // namespace detail

template <typename func_t, typename... args_t>
hipError_t cooperative(hipStream_t stream, const func_t& kernel,
                       std::size_t number_of_blocks,
                       std::size_t threads_per_block, args_t&&... args) {
  constexpr auto n = sizeof...(args_t) == 0 ? 1 : sizeof...(args_t);
  void* argument_ptrs[n];
  detail::for_each_argument_address(argument_ptrs, std::forward<args_t>(args)...);
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&kernel),
                                    dim3(number_of_blocks), dim3(threads_per_block),
                                    argument_ptrs, 0, stream);
}

template <typename func_t, typename... args_t>
void non_cooperative(hipStream_t stream, const func_t& kernel,
                     dim3 number_of_blocks, dim3 threads_per_block,
                     args_t&&... args) {
  kernel<<<number_of_blocks, threads_per_block, 0, stream>>>(
      std::forward<args_t>(args)...);
}

}  // namespace launch
}  // namespace loops
