/**
 * @file coo.hxx
 * @brief Coordinate-format matrix (reference include/loops/container/
 * coo.hxx:35-165), host side.  sort_by_row() orders entries by (row, col)
 * as the reference's zip-key sort does (coo.hxx:104-110); the conversion to
 * CSR offsets is the lower_bound of each row id (convert.hxx:69-78).
 */
#pragma once

#include <algorithm>
#include <cstddef>
#include <numeric>
#include <vector>

#include <loops/container/vector.hxx>

namespace loops {

template <typename index_t, typename value_t,
          memory_space_t space = memory_space_t::host>
struct coo_t {
  std::size_t rows = 0;
  std::size_t cols = 0;
  std::size_t nnzs = 0;
  vector_t<index_t, memory_space_t::host> row_indices;
  vector_t<index_t, memory_space_t::host> col_indices;
  vector_t<value_t, memory_space_t::host> values;

  coo_t() = default;
  coo_t(std::size_t r, std::size_t c, std::size_t nnz)
      : rows(r), cols(c), nnzs(nnz), row_indices(nnz), col_indices(nnz),
        values(nnz) {}

  void sort_by_row() {
    std::vector<std::size_t> perm(nnzs);
    std::iota(perm.begin(), perm.end(), std::size_t(0));
    std::stable_sort(perm.begin(), perm.end(),
                     [&](std::size_t a, std::size_t b) {
                       if (row_indices[a] != row_indices[b])
                         return row_indices[a] < row_indices[b];
                       return col_indices[a] < col_indices[b];
                     });
    vector_t<index_t, memory_space_t::host> r(nnzs), c(nnzs);
    vector_t<value_t, memory_space_t::host> v(nnzs);
    for (std::size_t i = 0; i < nnzs; ++i) {
      r[i] = row_indices[perm[i]];
      c[i] = col_indices[perm[i]];
      v[i] = values[perm[i]];
    }
    row_indices = std::move(r);
    col_indices = std::move(c);
    values = std::move(v);
  }
};

}  // namespace loops
