/**
 * @file search.hxx
 * @brief Merge-path diagonal search (reference include/loops/util/search.hxx:
 * 34-60).
 *
 * The merge path merges A = row-end offsets (a[i] = offsets[i+1], length
 * a_len = rows) with B = the counting sequence of nonzeros (length b_len =
 * nnz).  At diagonal d the reference returns x = the first idx in
 * [max(d-b_len,0), min(d,a_len)) with a[idx] > b[d-idx-1] (or the upper
 * bound), clamped to a_len, and y = d - x.  With f(k) = offsets[k] + k,
 * strictly increasing, that is x = max{k in [0, rows] : f(k) <= d}.
 *
 * Two implementations of the same function:
 *  - _binary_search: one work-item, ~log2(rows) dependent loads; used where
 *    every lane needs its own diagonal (work_oriented, per-thread LDS
 *    searches).
 *  - wave_search: a whole 64-lane wavefront cooperates on one diagonal,
 *    probing 64 evenly spaced candidates per round (ballot + popcount), so
 *    it needs ceil(log64(rows)) + 1 dependent loads instead of log2(rows).
 *    All 64 lanes must be active and call it with the same arguments.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <loops/container/coordinate.hxx>

namespace loops {
namespace search {

/// Sequential search: a is the row-end array (a[i] = end of row i), b a
/// counting array whose element k is b_base + k.
template <typename offset_t, typename a_t>
__host__ __device__ __forceinline__ coordinate_t<unsigned int> _binary_search(
    const offset_t& diagonal, const a_t* a, offset_t b_base,
    const offset_t& a_len, const offset_t& b_len) {
  long long d = static_cast<long long>(diagonal);
  long long x_min = d - static_cast<long long>(b_len);
  if (x_min < 0) x_min = 0;
  long long x_max = d < static_cast<long long>(a_len) ? d : a_len;
  long long first = x_min, count = x_max - x_min;
  while (count > 0) {
    long long step = count >> 1;
    long long idx = first + step;
    if (static_cast<long long>(a[idx]) <=
        static_cast<long long>(b_base) + (d - idx - 1)) {
      first = idx + 1;
      count -= step + 1;
    } else {
      count = step;
    }
  }
  long long x = first < static_cast<long long>(a_len) ? first : a_len;
  return coordinate_t<unsigned int>{static_cast<unsigned int>(x),
                                    static_cast<unsigned int>(d - first)};
}

/// Wavefront-cooperative search over the CSR offsets (rows + 1 entries).
/// Returns the same coordinate as _binary_search(d, offsets + 1, 0, rows,
/// nnz).  Must be called by all 64 lanes of a wave with uniform arguments.
__device__ __forceinline__ coordinate_t<unsigned int> wave_search(
    const int* __restrict__ offsets, int rows, int nnz, long long d) {
  const int lane = __lane_id();
  if (d > (long long)rows + nnz) d = (long long)rows + nnz;
  long long lo = d - nnz > 0 ? d - nnz : 0;    // f(lo) <= d always
  long long hi = d < rows ? d : rows;          // answer <= hi
  while (lo < hi) {
    long long span = hi - lo;
    long long step = (span + 63) >> 6;         // >= 1
    long long k = lo + (long long)(lane + 1) * step;
    bool pred = false;
    if (k <= hi) pred = (long long)offsets[k] + k <= d;
    // f is increasing, so the true lanes form a prefix; count them.
    int c = __popcll(__ballot(pred));
    long long new_hi = lo + (long long)(c + 1) * step - 1;
    lo = lo + (long long)c * step;
    hi = new_hi < hi ? new_hi : hi;
  }
  return coordinate_t<unsigned int>{static_cast<unsigned int>(lo),
                                    static_cast<unsigned int>(d - lo)};
}

}  // namespace search
}  // namespace loops
