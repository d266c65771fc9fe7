// SpMM building blocks shared by the single-matrix kernels and the batch
// kernel: column slices of B and C per lane, the merge-path unit walked by a
// group of lanes spanning the output columns, and its fixed-order fix-up.
#pragma once

#include <hip/hip_runtime.h>

#include <loops/util/search.hxx>

#include "device_common.hpp"
#include "loops_internal.hpp"

namespace loops_mi355x {

// ---------------------------------------------------------------------------
// Column slice of one lane: CPL consecutive columns starting at c.
// ---------------------------------------------------------------------------
template <int CPL>
struct cols_t {
  float v[CPL];
};

template <int CPL>
__device__ __forceinline__ cols_t<CPL> load_b(const float* __restrict__ B,
                                              long long ldb, int col, int c,
                                              int k) {
  cols_t<CPL> r;
  const float* p = B + (long long)col * ldb + c;
  if constexpr (CPL == 4) {
    if (c + 4 <= k) {
      f32x4 q = *reinterpret_cast<const f32x4*>(p);
      r.v[0] = q.x; r.v[1] = q.y; r.v[2] = q.z; r.v[3] = q.w;
      return r;
    }
  }
#pragma unroll
  for (int j = 0; j < CPL; ++j) r.v[j] = (c + j < k) ? p[j] : 0.f;
  return r;
}

template <int CPL>
__device__ __forceinline__ void store_c(float* __restrict__ p, int c, int k,
                                        const float* acc) {
  if constexpr (CPL == 4) {
    if (c + 4 <= k) {
      f32x4 q;
      q.x = acc[0]; q.y = acc[1]; q.z = acc[2]; q.w = acc[3];
      *reinterpret_cast<f32x4*>(p) = q;
      return;
    }
  }
#pragma unroll
  for (int j = 0; j < CPL; ++j)
    if (c + j < k) p[j] = acc[j];
}

// ---------------------------------------------------------------------------
// merge_path: group u walks merge diagonals [u*IPU, (u+1)*IPU).
// Work space per group: carry_row (int), carry[k] and head[k] (fp32).
// ---------------------------------------------------------------------------
struct spmm_ws {
  int* carry_row;
  float* carry;  // [units][k]
  float* head;   // [units][k]
};

inline size_t spmm_ws_bytes_impl(long long units, int k) {
  size_t b = (size_t)units * sizeof(int);
  b = (b + 255) & ~size_t(255);
  b += 2 * (size_t)units * k * sizeof(float);
  return (b + 255) & ~size_t(255);
}

inline spmm_ws carve_spmm_ws(void* ws, long long units, int k) {
  char* p = static_cast<char*>(ws);
  size_t b = ((size_t)units * sizeof(int) + 255) & ~size_t(255);
  spmm_ws w;
  w.carry_row = reinterpret_cast<int*>(p);
  w.carry = reinterpret_cast<float*>(p + b);
  w.head = w.carry + (size_t)units * k;
  return w;
}

// One merge-path unit u (merge diagonals [u*ipu, (u+1)*ipu)) walked by the
// calling group of G lanes (all G lanes call with the same u).
template <int G, int CPL, typename V>
__device__ __forceinline__ void spmm_merge_unit(
    const int* __restrict__ offsets, const int* __restrict__ indices,
    const V* __restrict__ values, int rows, int nnz, int k,
    const float* __restrict__ B, long long ldb, float* __restrict__ C,
    long long ldc, long long ipu, long long u, const spmm_ws& ws) {
  constexpr int U = 8;
  const int lane = threadIdx.x & (G - 1);
  const long long total = (long long)rows + nnz;
  long long d0 = u * ipu, d1 = d0 + ipu;
  if (d1 > total) d1 = total;
  const auto st = loops::search::_binary_search<long long>(
      d0, offsets + 1, 0LL, (long long)rows, (long long)nnz);
  const auto en = loops::search::_binary_search<long long>(
      d1, offsets + 1, 0LL, (long long)rows, (long long)nnz);
  const int xs = (int)st.x, ys = (int)st.y, xe = (int)en.x, ye = (int)en.y;
  const bool open_left = xs < xe && offsets[xs] < ys;

  for (int cb = 0; cb < k; cb += G * CPL) {
    const int c = cb + lane * CPL;
    int x = xs, y = ys;
    // Row ends of rows [rbase, rbase + G) held one per lane.
    int rbase = xs;
    int re_reg = offsets[(xs + 1 + lane) < rows ? (xs + 1 + lane) : rows];
    int row_end = __shfl(re_reg, 0, G);
    bool first = true;
    float acc[CPL];
#pragma unroll
    for (int j = 0; j < CPL; ++j) acc[j] = 0.f;

    auto emit = [&]() {
      if (first && open_left) {
        if (c < k) store_c<CPL>(ws.head + u * k + c, c, k, acc);
      } else if (c < k) {
        store_c<CPL>(C + (long long)x * ldc + c, c, k, acc);
      }
      first = false;
#pragma unroll
      for (int j = 0; j < CPL; ++j) acc[j] = 0.f;
      ++x;
      if (x - rbase >= G) {
        rbase = x;
        int r = x + 1 + lane;
        re_reg = offsets[r < rows ? r : rows];
      }
      row_end = __shfl(re_reg, x - rbase, G);
    };

    while (y < ye) {
      const int n = ye - y < U ? ye - y : U;
      int cj[U];
      float vj[U];
#pragma unroll
      for (int j = 0; j < U; ++j) {
        cj[j] = j < n ? indices[y + j] : 0;
        vj[j] = j < n ? value_traits<V>::load1(values + y + j) : 0.f;
      }
      cols_t<CPL> bj[U];
#pragma unroll
      for (int j = 0; j < U; ++j)
        if (j < n) bj[j] = load_b<CPL>(B, ldb, cj[j], c, k);
#pragma unroll
      for (int j = 0; j < U; ++j) {
        if (j < n) {
          while (x < xe && row_end <= y) emit();
#pragma unroll
          for (int q = 0; q < CPL; ++q) acc[q] += vj[j] * bj[j].v[q];
          ++y;
        }
      }
    }
    while (x < xe) emit();
    // acc now holds this group's part of row xe (if any).
    if (c < k) store_c<CPL>(ws.carry + u * k + c, c, k, acc);
  }
  if (lane == 0)
    ws.carry_row[u] = (xe < rows && offsets[xe] < ye) ? xe : -1;
}

template <int TPB, int G, int CPL, typename V>
__global__ __launch_bounds__(TPB) void spmm_merge_path_kernel(
    const int* __restrict__ offsets, const int* __restrict__ indices,
    const V* __restrict__ values, int rows, int nnz, int k,
    const float* __restrict__ B, long long ldb, float* __restrict__ C,
    long long ldc, long long ipu, long long units, spmm_ws ws) {
  const long long u = ((long long)blockIdx.x * TPB + threadIdx.x) / G;
  if (u >= units) return;  // uniform within the group
  spmm_merge_unit<G, CPL, V>(offsets, indices, values, rows, nnz, k, B, ldb, C,
                             ldc, ipu, u, ws);
}

// Per (unit, column): the unit where an open row starts sums the unit
// partials of that row in unit order (see spmv_fixup_kernel).
// Called by a whole workgroup for unit t; its lanes stride over the columns.
__device__ __forceinline__ void spmm_fixup_unit(const int* __restrict__ offsets,
                                                int rows, int k, long long ipu,
                                                long long t, const spmm_ws& ws,
                                                float* __restrict__ C,
                                                long long ldc) {
  const int r = ws.carry_row[t];
  if (r < 0 || r >= rows) return;
  const long long fr = (long long)offsets[r] + r;
  if (fr / ipu != t) return;
  const long long t_end = ((long long)offsets[r + 1] + r + 1 - 1) / ipu;
  for (int c = threadIdx.x; c < k; c += blockDim.x) {
    float s = ws.carry[t * k + c];
    for (long long v = t + 1; v < t_end; ++v) s += ws.carry[v * k + c];
    s += ws.head[t_end * k + c];
    C[(long long)r * ldc + c] = s;
  }
}



}  // namespace loops_mi355x
