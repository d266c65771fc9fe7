// gfx950 SpMM kernels, C = A * B with B (cols x k) and C (rows x k) dense
// row-major fp32 (leading dimensions ldb, ldc).
//
//   thread_mapped  algorithms/spmm/thread_mapped.cuh:28-53 (one row per
//                  work-item).  The reference rescans the row once per output
//                  column; here a work-item keeps 8 columns in registers and
//                  rescans once per 8 columns.  Every C(row, c) is still the
//                  sequential CSR-order fp32 sum of the reference.
#include <hip/hip_runtime.h>

#include <loops/schedule.hxx>

#include "device_common.hpp"
#include "loops_internal.hpp"

namespace loops_mi355x {

using loops::schedule::algorithms_t;

template <typename V>
__global__ void spmm_thread_mapped_kernel(
    loops::schedule::setup<algorithms_t::thread_mapped, 1, 1, int, int,
                           long long, long long> config,
    const int* __restrict__ indices, const V* __restrict__ values, int k,
    const float* __restrict__ B, long long ldb, float* __restrict__ C,
    long long ldc) {
  constexpr int CB = 8;
  for (auto row : config.tiles()) {
    for (int c0 = 0; c0 < k; c0 += CB) {
      float acc[CB];
#pragma unroll
      for (int j = 0; j < CB; ++j) acc[j] = 0.f;
      const int cn = k - c0 < CB ? k - c0 : CB;
      for (auto nz : config.atoms(row)) {
        const float v = value_traits<V>::load1(values + nz);
        const float* b = B + (long long)indices[nz] * ldb + c0;
#pragma unroll
        for (int j = 0; j < CB; ++j)
          if (j < cn) acc[j] += v * b[j];
      }
      float* c = C + (long long)row * ldc + c0;
#pragma unroll
      for (int j = 0; j < CB; ++j)
        if (j < cn) c[j] = acc[j];
    }
  }
}

namespace {
inline hipError_t record(hipEvent_t e, hipStream_t s) {
  return e ? hipEventRecord(e, s) : hipSuccess;
}

template <typename V>
hipError_t tm_spmm_t(const csr_args& A, int k, const float* B, long long ldb,
                     float* C, long long ldc, int tpb, hipEvent_t ev0,
                     hipEvent_t ev1, hipStream_t s) {
  hipError_t e = record(ev0, s);
  if (e != hipSuccess) return e;
  long long grid = ((long long)A.rows + tpb - 1) / tpb;
  if (grid > 0 && k > 0) {
    loops::schedule::setup<algorithms_t::thread_mapped, 1, 1, int, int,
                           long long, long long>
        config(const_cast<int*>(A.offsets), A.rows, A.nnz);
    spmm_thread_mapped_kernel<V><<<dim3((unsigned)grid), dim3(tpb), 0, s>>>(
        config, A.indices, static_cast<const V*>(A.values), k, B, ldb, C, ldc);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return record(ev1, s);
}
}  // namespace

hipError_t launch_tm_spmm(const csr_args& A, int k, const float* B, long long ldb,
                          float* C, long long ldc, int tpb, hipEvent_t ev0,
                          hipEvent_t ev1, hipStream_t s) {
  switch (A.vtype) {
    case LOOPS_VALUE_F32: return tm_spmm_t<float>(A, k, B, ldb, C, ldc, tpb, ev0, ev1, s);
    case LOOPS_VALUE_BF16: return tm_spmm_t<bf16_t>(A, k, B, ldb, C, ldc, tpb, ev0, ev1, s);
    case LOOPS_VALUE_F16: return tm_spmm_t<f16_t>(A, k, B, ldb, C, ldc, tpb, ev0, ev1, s);
  }
  return hipErrorInvalidValue;
}

bool gm_spmm_supported(int) { return false; }
hipError_t launch_gm_spmm(const csr_args&, int, const float*, long long, float*,
                          long long, int, hipEvent_t, hipEvent_t, hipStream_t) {
  return hipErrorNotSupported;
}
bool mp_spmm_supported(int, int) { return false; }
long long mp_spmm_units(const csr_args&, int, int) { return 0; }
size_t mp_spmm_ws_bytes(long long, int) { return 0; }
hipError_t launch_mp_spmm(const csr_args&, int, const float*, long long, float*,
                          long long, int, int, void*, hipEvent_t, hipEvent_t,
                          hipStream_t) {
  return hipErrorNotSupported;
}

}  // namespace loops_mi355x
