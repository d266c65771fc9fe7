/**
 * @file merge_path.cuh
 * @brief SpMM C = A B (B: cols x n, C: rows x n, dense row-major fp32,
 * container/matrix.cuh), merge_path schedule.
 * No counterpart in the reference: lane groups spanning the columns of B and C walk equal merge-path shares.  Blocks until done.
 */
#pragma once

#include <loops/algorithms/detail/capi.hxx>
#include <loops/container/matrix.cuh>

namespace loops {
namespace algorithms {
namespace spmm {

template <typename index_t, typename offset_t, typename type_t>
void merge_path(csr_t<index_t, offset_t, type_t>& csr, matrix_t<type_t>& B,
          matrix_t<type_t>& C, hipStream_t stream = 0) {
  detail::spmm(LOOPS_SCHEDULE_MERGE_PATH_FLAT, csr, B, C, stream);
}

}  // namespace spmm
}  // namespace algorithms
}  // namespace loops
