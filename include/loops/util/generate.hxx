/**
 * @file generate.hxx
 * @brief Random fills and random sparse matrices (reference include/loops/
 * util/generate.hxx:20-113).
 *
 * uniform_distribution(first, last, lo, hi, seed) fills [first, last) --
 * pointers into host or device memory -- with values uniform in [lo, hi)
 * for floating-point bounds and in [lo, hi] (both ends) for integer
 * bounds, as the reference's real / integer distributions are.  Element i
 * draws from its own seed, hash(i) mixed with `seed` (the reference's
 * per-element scheme, generate.hxx:45-61), so a fill is the same on the
 * host and on the device, and reproducible for a given seed; the default
 * seed is the clock, as in the reference.  The random stream itself is
 * this library's (the finaliser of MurmurHash3, a public-domain 32-bit
 * mixer, over the element's seed), not thrust's minstd engine.
 *
 * csr(rows, cols, sparsity, matrix) draws sparsity * rows * cols (row,
 * column) pairs and values in [0, 1), drops repeated pairs (the first
 * drawn is kept) and builds the CSR matrix (generate.hxx:85-113).  Row and
 * column ids are drawn in [0, rows - 1] and [0, cols - 1]; the reference
 * draws them in [0, rows] and [0, cols], which can produce an id one past
 * the end.  An optional fifth argument fixes the seed.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstddef>
#include <type_traits>

#include <loops/container/formats.hxx>
#include <loops/error.hxx>
#include <loops/memory.hxx>

namespace loops {
namespace generate {
namespace random {

/// 32-bit integer mixer (MurmurHash3's finaliser).
__host__ __device__ __forceinline__ unsigned int hash(unsigned int a) {
  a ^= a >> 16;
  a *= 0x85ebca6bu;
  a ^= a >> 13;
  a *= 0xc2b2ae35u;
  a ^= a >> 16;
  return a;
}

inline unsigned int clock_seed() {
  return (unsigned int)std::chrono::system_clock::now().time_since_epoch().count();
}

namespace detail {

// Value of element i for seed s, bounds of type bound_t.
template <typename bound_t>
__host__ __device__ __forceinline__ bound_t uniform_value(std::size_t i, unsigned int s,
                                                          bound_t lo, bound_t hi) {
  const unsigned int e = hash((unsigned int)i ^ hash((unsigned int)(i >> 32) + 0x9e3779b9u));
  const unsigned int bits = hash(e * 0x2545f491u + s);
  if constexpr (std::is_floating_point<bound_t>::value) {
    const double u = (double)(bits >> 8) * (1.0 / 16777216.0);  // [0, 1), 24 bits
    bound_t v = lo + (hi - lo) * (bound_t)u;
    return v < hi ? v : lo;  // a product rounded up to hi
  } else {
    const unsigned long long span =
        (unsigned long long)((long long)hi - (long long)lo) + 1ull;
    const unsigned long long pick = ((unsigned long long)bits * span) >> 32;
    return (bound_t)((long long)lo + (long long)pick);
  }
}

template <typename elem_t, typename bound_t>
__global__ void uniform_fill_kernel(elem_t* out, std::size_t n, bound_t lo, bound_t hi,
                                    unsigned int s) {
  for (std::size_t i = (std::size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (std::size_t)gridDim.x * blockDim.x)
    out[i] = (elem_t)uniform_value(i, s, lo, hi);
}

// True for device memory; host memory (registered or not), and any pointer
// when no GPU can be queried, is filled on the host.
inline bool is_device_pointer(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // clear the failed query
    return false;
  }
  return a.type == hipMemoryTypeDevice;
}

}  // namespace detail

template <typename iterator_t, typename type_t>
void uniform_distribution(iterator_t begin_it, iterator_t end_it, type_t begin, type_t end,
                          unsigned int useed = clock_seed()) {
  static_assert(std::is_pointer<iterator_t>::value,
                "uniform_distribution fills host or device arrays through pointers");
  using elem_t = typename std::remove_cv<typename std::remove_pointer<iterator_t>::type>::type;
  const std::size_t n = (std::size_t)(end_it - begin_it);
  if (n == 0) return;
  if (detail::is_device_pointer(begin_it)) {
    std::size_t grid = (n + 255) / 256;
    if (grid > 65536) grid = 65536;
    detail::uniform_fill_kernel<elem_t, type_t>
        <<<dim3((unsigned)grid), dim3(256), 0, 0>>>(begin_it, n, begin, end, useed);
    error::throw_if_exception(hipGetLastError(), "uniform_distribution launch");
    error::throw_if_exception(hipStreamSynchronize(0), "uniform_distribution");
  } else {
    for (std::size_t i = 0; i < n; ++i)
      begin_it[i] = (elem_t)detail::uniform_value(i, useed, begin, end);
  }
}

template <typename index_t, typename offset_t, typename value_t, memory_space_t space>
void csr(std::size_t rows, std::size_t cols, float sparsity,
         csr_t<index_t, offset_t, value_t, space>& matrix, unsigned int useed = clock_seed()) {
  const std::size_t nnzs =
      rows && cols && sparsity > 0.f ? (std::size_t)(sparsity * (rows * cols)) : 0;
  coo_t<index_t, value_t, memory_space_t::host> coo(rows, cols, nnzs);
  if (nnzs) {
    uniform_distribution(coo.row_indices.begin(), coo.row_indices.end(), index_t(0),
                         index_t(rows - 1), useed);
    uniform_distribution(coo.col_indices.begin(), coo.col_indices.end(), index_t(0),
                         index_t(cols - 1), hash(useed));
    uniform_distribution(coo.values.begin(), coo.values.end(), value_t(0), value_t(1),
                         hash(hash(useed)));
    coo.remove_duplicates();
  }
  matrix = csr_t<index_t, offset_t, value_t, space>(coo);
}

}  // namespace random
}  // namespace generate
}  // namespace loops
