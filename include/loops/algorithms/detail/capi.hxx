/**
 * @file capi.hxx
 * @brief Glue between the C++ drop-in entry points and the C ABI
 * (include/loops_capi.h): a per-thread library handle, status-to-exception
 * conversion and the non-owning CSR view of a csr_t.
 */
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <string>
#include <type_traits>

#include <loops/container/csr.hxx>
#include <loops/error.hxx>
#include <loops_capi.h>

namespace loops {
namespace algorithms {
namespace detail {

inline void check(loops_status_t s, const char* where) {
  if (s != LOOPS_SUCCESS) {
    std::string msg = std::string(where) + ": " + loops_status_string(s);
    if (s == LOOPS_ERROR_HIP)
      throw error::exception_t(static_cast<hipError_t>(loops_last_hip_error()), msg);
    throw error::exception_t(msg);
  }
}

/// The calling thread's handle (created on first use, destroyed at thread
/// exit).
inline loops_handle_t handle() {
  struct holder {
    loops_handle_t h = nullptr;
    ~holder() {
      if (h) loops_destroy(h);
    }
  };
  thread_local holder t;
  if (!t.h) check(loops_create(&t.h), "loops_create");
  return t.h;
}

/// The value type of type_t: float (loops_spmv), double or int
/// (loops_spmv_typed, the reference's other instantiations,
/// benchmarks/spmv/work_oriented.cu:53).
template <typename type_t>
constexpr loops_value_type_t value_type_of() {
  static_assert(std::is_same<type_t, float>::value || std::is_same<type_t, double>::value ||
                    (std::is_same<type_t, int>::value && sizeof(int) == 4),
                "the C++ entry points take float, double or int values (bf16 / fp16 "
                "storage is available through the C ABI's value_type)");
  return std::is_same<type_t, float>::value    ? LOOPS_VALUE_F32
         : std::is_same<type_t, double>::value ? LOOPS_VALUE_F64
                                               : LOOPS_VALUE_I32;
}

template <typename index_t, typename offset_t, typename type_t>
loops_csr_view_t view(const csr_t<index_t, offset_t, type_t, memory_space_t::device>& csr) {
  static_assert(std::is_integral<index_t>::value && sizeof(index_t) == 4,
                "the library takes 32-bit column indices");
  static_assert(std::is_integral<offset_t>::value && sizeof(offset_t) == 4,
                "the library takes 32-bit row offsets");
  loops_csr_view_t v;
  v.rows = (int64_t)csr.rows;
  v.cols = (int64_t)csr.cols;
  v.nnzs = (int64_t)csr.nnzs;
  v.offsets = reinterpret_cast<const int32_t*>(csr.offsets.raw());
  v.indices = reinterpret_cast<const int32_t*>(csr.indices.raw());
  v.values = csr.values.raw();
  v.value_type = value_type_of<type_t>();
  return v;
}

/// y = A x with `schedule`, synchronised on `stream` like the reference's
/// host entries; events (optional) are recorded around the main kernel.
template <typename index_t, typename offset_t, typename type_t>
void spmv(loops_schedule_t schedule,
          csr_t<index_t, offset_t, type_t, memory_space_t::device>& csr,
          vector_t<type_t, memory_space_t::device>& x,
          vector_t<type_t, memory_space_t::device>& y, hipStream_t stream,
          hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr) {
  error::throw_if_exception(x.size() < csr.cols, "x shorter than cols");
  error::throw_if_exception(y.size() < csr.rows, "y shorter than rows");
  loops_csr_view_t v = view(csr);
  loops_params_t p{};
  p.start_event = ev0;
  p.stop_event = ev1;
  if constexpr (std::is_same<type_t, float>::value)
    check(loops_spmv(handle(), schedule, &v, x.raw(), y.raw(), &p, nullptr, 0, stream),
          loops_schedule_name(schedule));
  else
    check(loops_spmv_typed(handle(), schedule, &v, x.raw(), y.raw(), &p, nullptr, 0, stream),
          loops_schedule_name(schedule));
  error::throw_if_exception(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

template <typename index_t, typename offset_t, typename type_t, typename matrix_t_>
void spmm(loops_schedule_t schedule,
          csr_t<index_t, offset_t, type_t, memory_space_t::device>& csr,
          matrix_t_& B, matrix_t_& C, hipStream_t stream) {
  error::throw_if_exception(B.rows < csr.cols, "B has fewer rows than A has columns");
  error::throw_if_exception(C.rows < csr.rows || C.cols < B.cols, "C too small");
  static_assert(std::is_same<type_t, float>::value, "SpMM takes fp32 values");
  loops_csr_view_t v = view(csr);
  loops_params_t p{};
  check(loops_spmm(handle(), schedule, &v, (int64_t)B.cols, B.m_data_ptr,
                   (int64_t)B.cols, C.m_data_ptr, (int64_t)C.cols, &p, nullptr,
                   0, stream),
        loops_schedule_name(schedule));
  error::throw_if_exception(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

}  // namespace detail
}  // namespace algorithms
}  // namespace loops
