/**
 * @file thread_mapped.cuh
 * @brief SpMV y = A x, thread_mapped schedule (one row per work-item through setup<thread_mapped>).
 * Drop-in for the reference's loops::algorithms::spmv::thread_mapped
 * (algorithms/spmv/thread_mapped.cuh:68-91): same signature, blocks until done.  Runs the library's gfx950
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
void thread_mapped(csr_t<index_t, offset_t, type_t>& csr, vector_t<type_t>& x,
          vector_t<type_t>& y, hipStream_t stream = 0) {
  detail::spmv(LOOPS_SCHEDULE_THREAD_MAPPED, csr, x, y, stream);
}

}  // namespace spmv
}  // namespace algorithms
}  // namespace loops
