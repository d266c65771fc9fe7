/**
 * @file sample.hxx
 * @brief The 4 x 4 sample matrix of the reference (include/loops/util/
 * sample.hxx:23-91), in host or device memory:
 *
 *   row 0: (empty)
 *   row 1: (1, 0) = 5, (1, 1) = 8
 *   row 2: (2, 2) = 3
 *   row 3: (3, 1) = 6
 *
 * i.e. offsets {0, 0, 2, 3, 4}, indices {0, 1, 2, 1}, values {5, 8, 3, 6};
 * with x = 1 the product is {0, 13, 3, 6}.
 */
#pragma once

#include <loops/container/formats.hxx>
#include <loops/memory.hxx>

namespace loops {
namespace sample {

template <memory_space_t space = memory_space_t::device, typename index_t = int,
          typename offset_t = int, typename value_t = float>
csr_t<index_t, offset_t, value_t, space> csr() {
  csr_t<index_t, offset_t, value_t, memory_space_t::host> m(4, 4, 4);
  const offset_t off[5] = {0, 0, 2, 3, 4};
  const index_t idx[4] = {0, 1, 2, 1};
  const value_t val[4] = {value_t(5), value_t(8), value_t(3), value_t(6)};
  for (int i = 0; i < 5; ++i) m.offsets[i] = off[i];
  for (int i = 0; i < 4; ++i) {
    m.indices[i] = idx[i];
    m.values[i] = val[i];
  }
  return csr_t<index_t, offset_t, value_t, space>(m);
}

}  // namespace sample
}  // namespace loops
