/**
 * @file csr.hxx
 * @brief Compressed sparse row matrix (reference include/loops/container/
 * csr.hxx:32-95): rows, cols, nnzs and the owned arrays offsets (Ap, rows+1),
 * indices (Aj, nnzs) and values (Ax, nnzs), in host or device memory.
 */
#pragma once

#include <cstddef>

#include <loops/container/coo.hxx>
#include <loops/container/vector.hxx>

namespace loops {

template <typename index_t, typename offset_t, typename value_t,
          memory_space_t space = memory_space_t::device>
struct csr_t {
  std::size_t rows;
  std::size_t cols;
  std::size_t nnzs;

  vector_t<offset_t, space> offsets;  /// Ap
  vector_t<index_t, space> indices;   /// Aj
  vector_t<value_t, space> values;    /// Ax

  csr_t() : rows(0), cols(0), nnzs(0) {}

  csr_t(std::size_t r, std::size_t c, std::size_t nnz)
      : rows(r), cols(c), nnzs(nnz), offsets(r + 1), indices(nnz),
        values(nnz) {}

  /// Copy between memory spaces.
  template <memory_space_t rhs_space>
  csr_t(const csr_t<index_t, offset_t, value_t, rhs_space>& rhs)
      : rows(rhs.rows), cols(rhs.cols), nnzs(rhs.nnzs), offsets(rhs.offsets),
        indices(rhs.indices), values(rhs.values) {}

  /// From a (host) COO matrix: sort by (row, col), offsets by lower_bound.
  template <memory_space_t rhs_space>
  csr_t(const coo_t<index_t, value_t, rhs_space>& coo)
      : rows(coo.rows), cols(coo.cols), nnzs(coo.nnzs) {
    coo_t<index_t, value_t, memory_space_t::host> c(coo);
    c.sort_by_row();
    vector_t<offset_t, memory_space_t::host> off(rows + 1);
    detail::indices_to_offsets(c.row_indices, off);
    offsets = vector_t<offset_t, space>(off);
    indices = vector_t<index_t, space>(c.col_indices);
    values = vector_t<value_t, space>(c.values);
  }
};

}  // namespace loops
