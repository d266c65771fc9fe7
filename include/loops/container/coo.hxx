/**
 * @file coo.hxx
 * @brief Coordinate-format matrix (reference include/loops/container/
 * coo.hxx:35-165), host side.  sort_by_row() orders entries by (row, col)
 * and sort_by_column() by (column, row), as the reference's zip-key sorts do
 * (coo.hxx:104-122); both are stable, so entries with equal keys keep their
 * order (thrust's sort_by_key on those keys is a merge sort).
 * remove_duplicates() sorts by row and keeps the first entry of each
 * (row, column) pair (coo.hxx:127-145, unique_by_key).  A COO matrix can be
 * built from a CSR one: row ids from the offsets (convert.hxx:36-59), the
 * column ids and values copied.  (The reference's constructor reads a
 * `col_indices` member that its csr_t does not have, so it does not
 * compile when used; this one reads `indices`.)  The conversion to CSR
 * offsets is the lower_bound of each row id (convert.hxx:69-78).
 */
#pragma once

#include <algorithm>
#include <cstddef>
#include <numeric>
#include <type_traits>
#include <vector>

#include <loops/container/detail/convert.hxx>
#include <loops/container/vector.hxx>

namespace loops {

template <typename index_t, typename offset_t, typename value_t, memory_space_t space>
struct csr_t;

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

  /// From a COO matrix tagged with another memory space (the entries live
  /// in host memory either way).
  template <memory_space_t rhs_space,
            typename = typename std::enable_if<rhs_space != space>::type>
  coo_t(const coo_t<index_t, value_t, rhs_space>& rhs)
      : rows(rhs.rows), cols(rhs.cols), nnzs(rhs.nnzs), row_indices(rhs.row_indices),
        col_indices(rhs.col_indices), values(rhs.values) {}

  /// From a CSR matrix of either memory space.
  template <typename offset_t, memory_space_t rhs_space>
  coo_t(const csr_t<index_t, offset_t, value_t, rhs_space>& csr)
      : rows(csr.rows), cols(csr.cols), nnzs(csr.nnzs), row_indices(csr.nnzs),
        col_indices(csr.indices), values(csr.values) {
    detail::offsets_to_indices(csr.offsets, row_indices);
  }

  void sort_by_row() { sort_by(row_indices, col_indices); }
  void sort_by_column() { sort_by(col_indices, row_indices); }

  void remove_duplicates() {
    sort_by_row();
    std::size_t kept = 0;
    for (std::size_t i = 0; i < nnzs; ++i) {
      if (kept > 0 && row_indices[kept - 1] == row_indices[i] &&
          col_indices[kept - 1] == col_indices[i])
        continue;
      row_indices[kept] = row_indices[i];
      col_indices[kept] = col_indices[i];
      values[kept] = values[i];
      ++kept;
    }
    nnzs = kept;
    row_indices.resize(kept);
    col_indices.resize(kept);
    values.resize(kept);
  }

 private:
  // Stable sort of the entries by (major, minor).
  void sort_by(const vector_t<index_t, memory_space_t::host>& major,
               const vector_t<index_t, memory_space_t::host>& minor) {
    std::vector<std::size_t> perm(nnzs);
    std::iota(perm.begin(), perm.end(), std::size_t(0));
    std::stable_sort(perm.begin(), perm.end(), [&](std::size_t a, std::size_t b) {
      if (major[a] != major[b]) return major[a] < major[b];
      return minor[a] < minor[b];
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
