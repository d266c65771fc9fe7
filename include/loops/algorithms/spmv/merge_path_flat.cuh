/**
 * @file merge_path_flat.cuh
 * @brief SpMV y = A x, merge-path schedule: equal shares of (rows +
 * nonzeros) per workgroup and work-item, rows split across tiles finished by
 * a fixed-order fix-up.  Drop-in for the reference's
 * loops::algorithms::spmv::merge_path_flat (algorithms/spmv/
 * merge_path_flat.cuh:96-139): same signature, blocks until done and returns
 * a timer holding the merge-path kernel's time.
 */
#pragma once

#include <loops/algorithms/detail/capi.hxx>
#include <loops/container/vector.hxx>
#include <loops/util/timer.hxx>

namespace loops {
namespace algorithms {
namespace spmv {

template <typename index_t, typename offset_t, typename type_t>
util::timer_t merge_path_flat(csr_t<index_t, offset_t, type_t>& csr,
                              vector_t<type_t>& x, vector_t<type_t>& y,
                              hipStream_t stream = 0) {
  util::timer_t timer(stream);
  detail::spmv(LOOPS_SCHEDULE_MERGE_PATH_FLAT, csr, x, y, stream,
               timer.start_event(), timer.stop_event());
  (void)hipEventElapsedTime(&timer.time, timer.start_event(), timer.stop_event());
  return timer;
}

}  // namespace spmv
}  // namespace algorithms
}  // namespace loops
