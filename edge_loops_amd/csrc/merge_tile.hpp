// The merge-path tile (one workgroup, TPB lanes x IPT merge steps) and the
// fixed-order sums of rows that cross tiles, shared by the plain merge-path
// kernel and the column-panel kernel.  See spmv_kernels.hip for the scheme.
#pragma once

#include <hip/hip_runtime.h>

#include <loops/util/search.hxx>

#include "device_common.hpp"
#include "loops_internal.hpp"

namespace loops_mi355x {

// Sum of carry[t0 .. t1-1] + head[t1] by one wavefront, in an order fixed by
// (t1 - t0) alone: lane l adds elements l, l+64, ... in sequence, then the
// 64 lane sums are combined by a fixed butterfly.  All 64 lanes must call.
__device__ __forceinline__ float wave_segment_sum(const float* __restrict__ carry,
                                                  const float* __restrict__ head,
                                                  long long t0, long long t1) {
  const int lane = threadIdx.x & 63;
  float s = 0.f;
  for (long long u = t0 + lane; u < t1; u += 64) s += carry[u];
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
  return s + head[t1];
}

// The value of an open row: one work-item for a two-unit row (its one carry
// plus the head), the whole wavefront otherwise.  Both paths give the same
// bits for a two-unit row (the butterfly only adds exact zeros to it), so a
// row's value never depends on which path computed it.
__device__ __forceinline__ bool segment_is_short(long long t0, long long t1) {
  return t1 - t0 <= 1;
}
__device__ __forceinline__ float short_segment_sum(const float* __restrict__ carry,
                                                   const float* __restrict__ head,
                                                   long long t0, long long t1) {
  return (t0 < t1 ? carry[t0] : 0.f) + head[t1];
}

// ---------------------------------------------------------------------------
// Merge-path SpMV.
//
// One workgroup per merge tile of TI = TPB * IPT steps (rows + nonzeros).
//  1. Tile end points: read from a precomputed partition, or found by two
//     wavefronts, each running a 64-way cooperative search
//     (loops::search::wave_search).
//  2. The row ends of the tile's rows go to LDS.  The tile's nonzeros are
//     streamed with 16-byte non-temporal loads (column ids and values, all
//     issued before any x gather), and the products a_ij * x_j go to LDS, so
//     A is read once, coalesced, and x keeps the L2.
//  3. Each lane finds its start with a binary search in LDS and walks IPT
//     merge steps (consume a product, or close a row and store y).  A lane
//     keeps its first closed row back.
//  4. A workgroup segmented scan gives each lane the partial sum carried in
//     from earlier lanes; it is added to the lane's first closed row.
//  5. The row entering the tile already open and the row leaving it open go
//     to the fix-up slots.
// mp_tile() is the whole tile; the plain kernel runs one tile per workgroup,
// the column-panel kernel one tile of one panel per workgroup.
// ---------------------------------------------------------------------------
template <int TPB, int IPT>
struct mp_smem {
  static constexpr int TI = TPB * IPT;
  int rowend[TI + 1];
  __align__(16) float prod[TI + 4];
  int coord[5];
  float wv[TPB / 64];
  int wf[TPB / 64];
};

template <int TPB, int IPT, typename V, bool VEC>
__device__ __forceinline__ void mp_tile(
    const int* __restrict__ offsets, const int* __restrict__ indices,
    const V* __restrict__ values, const float* __restrict__ x,
    float* __restrict__ y, int rows, int nnz, int tile,
    const loops_coord_t* __restrict__ coords, const fixup_ws& ws,
    mp_smem<TPB, IPT>& sm) {
  constexpr int TI = TPB * IPT;
  constexpr int NC = (TI + 4 * TPB - 1) / (4 * TPB);  // 4-wide chunks per lane
  constexpr int NW = TPB / 64;
  int* s_rowend = sm.rowend;
  float* s_prod = sm.prod;
  int* s_coord = sm.coord;
  float* s_wv = sm.wv;
  int* s_wf = sm.wf;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const long long total = (long long)rows + nnz;
  long long d0 = (long long)tile * TI;
  if (d0 > total) d0 = total;
  long long d1 = d0 + TI;
  if (d1 > total) d1 = total;

  // 1. tile end points
  if (coords != nullptr) {
    if (tid < 2) {
      loops_coord_t c = coords[tile + tid];
      s_coord[2 * tid] = (int)c.x;
      s_coord[2 * tid + 1] = (int)c.y;
    }
  } else {
    if (wave == 0) {
      auto c = loops::search::wave_search(offsets, rows, nnz, d0);
      if (lane == 0) {
        s_coord[0] = (int)c.x;
        s_coord[1] = (int)c.y;
      }
    }
    if (wave == (NW > 1 ? 1 : 0)) {
      auto c = loops::search::wave_search(offsets, rows, nnz, d1);
      if (lane == 0) {
        s_coord[2] = (int)c.x;
        s_coord[3] = (int)c.y;
      }
    }
  }
  __syncthreads();
  const int xs = s_coord[0], ys = s_coord[1];
  const int xe = s_coord[2], ye = s_coord[3];
  const int tile_rows = xe - xs;
  const int tile_nnz = ye - ys;
  const int tile_items = (int)(d1 - d0);

  // 2. Column ids and values of the tile's nonzeros [ys, ye) are loaded
  // first, the row ends next (both in flight together), then the x gathers
  // and the LDS stores of the products.
  const int ybase = ys & ~3;
  const int pofs = ys - ybase;
  {
    i32x4 ci[NC];
    f32x4 cv[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const int j = ybase + 4 * (tid + q * TPB);
      if (VEC && j >= ys && j + 4 <= ye) {
        ci[q] = ld_idx4(indices + j);
        cv[q] = value_traits<V>::load4(values + j);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int jj = j + e;
          const bool ok = jj >= ys && jj < ye;
          ci[q][e] = ok ? ld_idx(indices + jj) : 0;
          cv[q][e] = ok ? value_traits<V>::load1(values + jj) : 0.f;
        }
      }
    }
    // 2a. row ends of rows xs .. xe, relative to ys; plus the start of row xs
    for (int i = tid; i <= tile_rows; i += TPB) {
      int r = xs + 1 + i;
      if (r > rows) r = rows;
      s_rowend[i] = offsets[r] - ys;
    }
    if (tid == 0) s_coord[4] = offsets[xs] - ys;

#pragma unroll
    for (int q = 0; q < NC; ++q) {
      const int j = ybase + 4 * (tid + q * TPB);
      if (j < ye) {
        f32x4 p;
        p.x = cv[q].x * x[ci[q].x];
        p.y = cv[q].y * x[ci[q].y];
        p.z = cv[q].z * x[ci[q].z];
        p.w = cv[q].w * x[ci[q].w];
        if (j >= ys && j + 4 <= ye) {
          *reinterpret_cast<f32x4*>(&s_prod[j - ybase]) = p;
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (j + e >= ys && j + e < ye) s_prod[j + e - ybase] = p[e];
        }
      }
    }
    // The chunks cover [ybase, ybase + 4 * TPB * NC), at least TI entries;
    // a tile of TI nonzeros starting off a 4-boundary ends up to 3 beyond.
    const int jt = ybase + 4 * TPB * NC + tid;
    if (jt < ye)
      s_prod[jt - ybase] = value_traits<V>::load1(values + jt) * x[ld_idx(indices + jt)];
  }
  __syncthreads();

  // 3. lane start by search in LDS, then IPT merge steps
  const int dl = tid * IPT;
  int steps = tile_items - dl;
  steps = steps < 0 ? 0 : (steps > IPT ? IPT : steps);
  int xl, yl;
  {
    const int dd = dl < tile_items ? dl : tile_items;
    int x_min = dd - tile_nnz;
    x_min = x_min < 0 ? 0 : x_min;
    const int x_max = dd < tile_rows ? dd : tile_rows;
    int first = x_min, count = x_max - x_min;
    while (count > 0) {
      int step = count >> 1;
      int idx = first + step;
      if (s_rowend[idx] <= dd - idx - 1) {
        first = idx + 1;
        count -= step + 1;
      } else {
        count = step;
      }
    }
    xl = first;
    yl = dd - first;
  }
  float run = 0.f;
  int first_row = -1;
  float first_val = 0.f;
#pragma unroll
  for (int s = 0; s < IPT; ++s) {
    if (s < steps) {
      if (yl < s_rowend[xl]) {
        run += s_prod[yl + pofs];
        ++yl;
      } else {
        if (first_row < 0) {
          first_row = xl;
          first_val = run;
        } else {
          y[xs + xl] = run;
        }
        run = 0.f;
        ++xl;
      }
    }
  }

  // 4. partial sum carried in from earlier lanes
  float v = run;
  int f = first_row >= 0 ? 1 : 0;
  group_seg_scan<64>(v, f);
  if (lane == 63) {
    s_wv[wave] = v;
    s_wf[wave] = f;
  }
  __syncthreads();
  float pv = 0.f;
  for (int w = 0; w < wave; ++w) {
    float wv = s_wv[w];
    pv = s_wf[w] ? wv : pv + wv;
  }
  float ev = __shfl_up(v, 1, 64);
  int ef = __shfl_up(f, 1, 64);
  const float excl = lane == 0 ? pv : (ef ? ev : pv + ev);
  const bool open_left = tile_rows > 0 && s_coord[4] < 0;
  if (first_row >= 0) {
    const float val = excl + first_val;
    if (first_row == 0 && open_left) {
      ws.head_val[tile] = val;
    } else {
      y[xs + first_row] = val;
    }
  }

  // 5. the row leaving the tile open
  const int xe_start = tile_rows > 0 ? s_rowend[tile_rows - 1] : s_coord[4];
  const bool open_right = xe < rows && xe_start < tile_nnz;
  if (tid == TPB - 1) {
    const float tot = f ? v : pv + v;
    ws.carry_row[tile] = open_right ? xe : -1;
    ws.carry_val[tile] = tot;
    int seg_end = -1;
    if (open_right) {
      const long long fr = (long long)xe_start + ys + xe;
      if (fr / TI == tile)
        seg_end = (int)(((long long)s_rowend[tile_rows] + ys + xe + 1 - 1) / TI);
    }
    ws.seg_end[tile] = seg_end;
  }
}

}  // namespace loops_mi355x
