/**
 * @file convert.hxx
 * @brief Offsets <-> indices conversions of the sparse containers
 * (reference include/loops/container/detail/convert.hxx:36-78).
 *
 * offsets_to_indices: entry k of a compressed (CSR / CSC) matrix gets the
 *   id of the row (column) holding it.  As in the reference, the id of each
 *   non-empty row is placed at its first entry, every other position starts
 *   at 0, and a running maximum fills the rest (convert.hxx:36-59); for
 *   well-formed offsets (non-decreasing, offsets[0] = 0) that is the row of
 *   every entry.
 * indices_to_offsets: offsets[i] = the first position whose (sorted) id is
 *   >= i, for every i < offsets.size() (convert.hxx:69-78, a lower_bound per
 *   id).
 *
 * These run on the host: device vectors are copied to host memory and the
 * result back (they are set-up steps, not on the SpMV path).
 */
#pragma once

#include <algorithm>
#include <cstddef>

#include <loops/container/vector.hxx>

namespace loops {
namespace detail {

template <typename offset_t, typename index_t>
void offsets_to_indices(const offset_t* offsets, std::size_t num_offsets,
                        index_t* indices, std::size_t num_indices) {
  std::fill(indices, indices + num_indices, index_t(0));
  for (std::size_t r = 0; r + 1 < num_offsets; ++r) {
    const offset_t at = offsets[r];
    if (at != offsets[r + 1] && at >= offset_t(0) && std::size_t(at) < num_indices)
      indices[at] = index_t(r);
  }
  index_t running = index_t(0);
  for (std::size_t k = 0; k < num_indices; ++k) {
    running = std::max(running, indices[k]);
    indices[k] = running;
  }
}

template <typename index_t, typename offset_t>
void indices_to_offsets(const index_t* indices, std::size_t num_indices,
                        offset_t* offsets, std::size_t num_offsets) {
  for (std::size_t i = 0; i < num_offsets; ++i)
    offsets[i] = offset_t(std::lower_bound(indices, indices + num_indices, index_t(i)) - indices);
}

/// Vectors of either memory space; `indices` keeps its size.
template <typename index_v_t, typename offset_v_t>
void offsets_to_indices(const offset_v_t& offsets, index_v_t& indices) {
  using offset_t = typename offset_v_t::value_type;
  using index_t = typename index_v_t::value_type;
  const vector_t<offset_t, memory_space_t::host> off(offsets);
  vector_t<index_t, memory_space_t::host> idx(indices.size());
  offsets_to_indices(off.data(), off.size(), idx.data(), idx.size());
  indices = index_v_t(idx);
}

/// Vectors of either memory space; `offsets` keeps its size.
template <typename index_v_t, typename offset_v_t>
void indices_to_offsets(const index_v_t& indices, offset_v_t& offsets) {
  using offset_t = typename offset_v_t::value_type;
  using index_t = typename index_v_t::value_type;
  const vector_t<index_t, memory_space_t::host> idx(indices);
  vector_t<offset_t, memory_space_t::host> off(offsets.size());
  indices_to_offsets(idx.data(), idx.size(), off.data(), off.size());
  offsets = offset_v_t(off);
}

}  // namespace detail
}  // namespace loops
