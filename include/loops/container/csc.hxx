/**
 * @file csc.hxx
 * @brief Compressed sparse column matrix (reference include/loops/container/
 * csc.hxx:21-91): rows, cols, nnzs and the owned arrays offsets (one per
 * column, plus one), indices (the row of each entry) and values, in host or
 * device memory.  Built from a COO matrix by sorting it by (column, row)
 * and taking the lower_bound of each column id as the column's first entry.
 *
 * The (rows, cols, nnzs) constructor sizes offsets to cols + 1 (the
 * reference's sizes them by rows + 1 there, csc.hxx:44-50, and by
 * cols + 1 when built from COO).
 */
#pragma once

#include <cstddef>

#include <loops/container/coo.hxx>
#include <loops/container/detail/convert.hxx>
#include <loops/container/vector.hxx>

namespace loops {

template <typename index_t, typename offset_t, typename value_t,
          memory_space_t space = memory_space_t::device>
struct csc_t {
  std::size_t rows;
  std::size_t cols;
  std::size_t nnzs;

  vector_t<offset_t, space> offsets;  /// Ap, cols + 1
  vector_t<index_t, space> indices;   /// Ai, the row of each entry
  vector_t<value_t, space> values;    /// Ax

  csc_t() : rows(0), cols(0), nnzs(0) {}

  csc_t(std::size_t r, std::size_t c, std::size_t nnz)
      : rows(r), cols(c), nnzs(nnz), offsets(c + 1), indices(nnz), values(nnz) {}

  /// Copy between memory spaces.
  template <memory_space_t rhs_space>
  csc_t(const csc_t<index_t, offset_t, value_t, rhs_space>& rhs)
      : rows(rhs.rows), cols(rhs.cols), nnzs(rhs.nnzs), offsets(rhs.offsets),
        indices(rhs.indices), values(rhs.values) {}

  /// From a COO matrix: sort by (column, row), offsets by lower_bound.
  template <memory_space_t rhs_space>
  csc_t(const coo_t<index_t, value_t, rhs_space>& coo)
      : rows(coo.rows), cols(coo.cols), nnzs(coo.nnzs) {
    coo_t<index_t, value_t, memory_space_t::host> c(coo);
    c.sort_by_column();
    vector_t<offset_t, memory_space_t::host> off(cols + 1);
    detail::indices_to_offsets(c.col_indices, off);
    offsets = vector_t<offset_t, space>(off);
    indices = vector_t<index_t, space>(c.row_indices);
    values = vector_t<value_t, space>(c.values);
  }
};

}  // namespace loops
