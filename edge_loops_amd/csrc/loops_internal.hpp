// Internal launch interface between the C ABI (capi.hip) and the kernel
// translation units.  Host-only declarations; every function launches on the
// given stream and returns the hipError_t of the launch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "loops_capi.h"

namespace loops_mi355x {

struct csr_args {
  const int* offsets;
  const int* indices;
  const void* values;
  loops_value_type_t vtype;
  int rows;
  int cols;
  int nnz;
  bool aligned;  // indices / values allow 16-byte (8-byte for 16-bit) loads
};

// Slots of the row fix-up: per work unit (merge tile or work_oriented
// thread) the row continuing past its end and two partial sums (caller
// workspace, no initialisation needed).  counter: per unit, used only by the
// in-kernel fix-up; it lives in memory owned by the handle that holds
// nothing else, is zero before a call and is left zero by it (the last
// participant of each row resets its counter).
// seg_end[u]: for the unit where its open row starts, the unit holding that
// row's end (so the fix-up needs no offsets lookup); -1 otherwise.
struct fixup_ws {
  int* carry_row;
  int* seg_end;
  float* carry_val;
  float* head_val;
  int* counter;
};

inline size_t fixup_ws_bytes(long long units) {
  size_t b = (size_t)units * (2 * sizeof(int) + 2 * sizeof(float));
  return (b + 255) & ~size_t(255);
}
fixup_ws carve_fixup_ws(void* ws, long long units, int* counters);

// ---- SpMV ----------------------------------------------------------------
bool mp_spmv_supported(int tpb, int ipt);
hipError_t launch_mp_spmv(const csr_args& A, const float* x, float* y, int tpb,
                          int ipt, const loops_coord_t* coords,
                          const fixup_ws& ws, bool fused, hipEvent_t ev0,
                          hipEvent_t ev1, hipStream_t s);
hipError_t launch_mp_partition(const int* offsets, int rows, int nnz,
                               long long items_per_tile, long long num_coords,
                               loops_coord_t* coords, hipStream_t s);
hipError_t launch_wo_partition(const int* offsets, int rows, int nnz,
                               long long num_threads, loops_coord_t* coords,
                               hipStream_t s);
hipError_t launch_fixup(const int* offsets, int rows, long long unit_items,
                        long long num_units, const fixup_ws& ws, float* y,
                        hipStream_t s);

bool wo_spmv_supported(int tpb);
hipError_t launch_wo_spmv(const csr_args& A, const float* x, float* y, int tpb,
                          int grid, const fixup_ws& ws, hipEvent_t ev0,
                          hipEvent_t ev1, hipStream_t s);

bool gm_spmv_supported(int tpb, int tpt);
hipError_t launch_gm_spmv(const csr_args& A, const float* x, float* y, int tpb,
                          int tpt, hipEvent_t ev0, hipEvent_t ev1,
                          hipStream_t s);

hipError_t launch_tm_spmv(const csr_args& A, const float* x, float* y, int tpb,
                          bool original, hipEvent_t ev0, hipEvent_t ev1,
                          hipStream_t s);

// ---- SpMM ----------------------------------------------------------------
hipError_t launch_tm_spmm(const csr_args& A, int k, const float* B, long long ldb,
                          float* C, long long ldc, int tpb, hipEvent_t ev0,
                          hipEvent_t ev1, hipStream_t s);
bool gm_spmm_supported(int tpb);
hipError_t launch_gm_spmm(const csr_args& A, int k, const float* B, long long ldb,
                          float* C, long long ldc, int tpb, hipEvent_t ev0,
                          hipEvent_t ev1, hipStream_t s);
bool mp_spmm_supported(int tpb, int ipw);
long long mp_spmm_units(const csr_args& A, int tpb, int ipw);
size_t mp_spmm_ws_bytes(long long units, int k);
hipError_t launch_mp_spmm(const csr_args& A, int k, const float* B, long long ldb,
                          float* C, long long ldc, int tpb, int ipw, void* ws,
                          hipEvent_t ev0, hipEvent_t ev1, hipStream_t s);

}  // namespace loops_mi355x
