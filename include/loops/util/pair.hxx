/**
 * @file pair.hxx
 * @brief Minimal device-usable pair and counting iterator (the schedules
 * return these where the reference used thrust::pair and
 * thrust::counting_iterator).
 */
#pragma once

#include <hip/hip_runtime.h>

namespace loops {

template <typename A, typename B>
struct pair_t {
  A first;
  B second;
};

template <typename A, typename B>
__host__ __device__ __forceinline__ pair_t<A, B> make_pair(A a, B b) {
  return pair_t<A, B>{a, b};
}

/// it[i] == base + i.
template <typename T>
struct counting_iterator {
  T base;
  __host__ __device__ counting_iterator(T b = T(0)) : base(b) {}
  __host__ __device__ T operator[](long long i) const { return base + T(i); }
  __host__ __device__ T operator*() const { return base; }
};

}  // namespace loops
