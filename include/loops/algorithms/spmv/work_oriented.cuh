/**
 * @file work_oriented.cuh
 * @brief SpMV y = A x, work_oriented schedule (an equal merge-path share per work-item, grid 2 x CU count).
 * Drop-in for the reference's loops::algorithms::spmv::work_oriented
 * (algorithms/spmv/work_oriented.cuh:91-113): same signature, blocks until done.  Runs the library's gfx950
 * kernel through the C ABI; y is overwritten (the reference's atomic
 * schedules add into a zeroed y).
 */
#pragma once

#include <loops/algorithms/detail/capi.hxx>
#include <loops/container/vector.hxx>

namespace loops {
namespace algorithms {
namespace spmv {

template <typename index_t, typename offset_t, typename type_t>
void work_oriented(csr_t<index_t, offset_t, type_t>& csr, vector_t<type_t>& x,
          vector_t<type_t>& y, hipStream_t stream = 0) {
  detail::spmv(LOOPS_SCHEDULE_WORK_ORIENTED, csr, x, y, stream);
}

}  // namespace spmv
}  // namespace algorithms
}  // namespace loops
