// gfx950 SpMV kernels for the five schedules, plus the merge-path partition
// and the deterministic row fix-up.
//
// Reference kernels restated here (paths relative to the reference root):
//   merge_path_flat  algorithms/spmv/merge_path_flat.cuh:38-83
//   work_oriented    algorithms/spmv/work_oriented.cuh:34-78
//   group_mapped     algorithms/spmv/group_mapped.cuh:26-60
//   thread_mapped    algorithms/spmv/thread_mapped.cuh:26-55
//   original         algorithms/spmv/original.cuh:26-46
//
// None of them uses floating-point atomics.  A row that crosses the end of a
// work unit (a merge tile, or a work_oriented thread's share) is finished by
// the fix-up: every unit records the row it leaves open (carry_row), its
// partial sum of that row (carry_val) and its partial sum of the row it
// entered already open (head_val).  Row r occupies merge diagonals
// [f(r), f(r+1)) with f(k) = offsets[k] + k, so the units holding it are
// t0 = f(r) / U .. t1 = (f(r+1) - 1) / U (U = merge steps per unit); its
// value is carry_val[t0] + ... + carry_val[t1 - 1] + head_val[t1], summed in
// that order, so every y[row] comes from one fixed sequence of fp32
// additions on every run.  The fix-up runs either as a second kernel, or
// inside the merge-path kernel: each participating unit bumps a counter
// (agent-scope acquire/release atomics) and the last of the n = t1 - t0 + 1
// participants computes the sum and resets the counter to zero.
#include <hip/hip_runtime.h>

#include <loops/schedule.hxx>
#include <loops/util/search.hxx>

#include "device_common.hpp"
#include "loops_internal.hpp"

namespace loops_mi355x {

using loops::schedule::algorithms_t;

fixup_ws carve_fixup_ws(void* ws, long long units, int* counters) {
  char* p = static_cast<char*>(ws);
  fixup_ws w;
  w.carry_row = reinterpret_cast<int*>(p);
  w.seg_end = reinterpret_cast<int*>(p + units * sizeof(int));
  w.carry_val = reinterpret_cast<float*>(p + units * 2 * sizeof(int));
  w.head_val = reinterpret_cast<float*>(p + units * (2 * sizeof(int) + sizeof(float)));
  w.counter = counters;
  return w;
}

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
//     through the fix-up (slots + optional in-kernel last-arriver).
// ---------------------------------------------------------------------------
template <int TPB, int IPT, typename V, bool VEC, bool FUSED>
__global__ __launch_bounds__(TPB) void spmv_merge_path_kernel(
    const int* __restrict__ offsets, const int* __restrict__ indices,
    const V* __restrict__ values, const float* __restrict__ x,
    float* __restrict__ y, int rows, int nnz,
    const loops_coord_t* __restrict__ coords, fixup_ws ws) {
  constexpr int TI = TPB * IPT;
  constexpr int NW = TPB / 64;
  constexpr int NC = (TI + 4 * TPB - 1) / (4 * TPB);  // 4-wide chunks per lane
  __shared__ int s_rowend[TI + 1];
  __shared__ __align__(16) float s_prod[TI + 4];
  __shared__ int s_coord[5];
  __shared__ float s_wv[NW];
  __shared__ int s_wf[NW];
  __shared__ float s_part[2];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int tile = blockIdx.x;
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

  // 2a. row ends of rows xs .. xe, relative to ys; plus the start of row xs
  for (int i = tid; i <= tile_rows; i += TPB) {
    int r = xs + 1 + i;
    if (r > rows) r = rows;
    s_rowend[i] = offsets[r] - ys;
  }
  if (tid == 0) s_coord[4] = offsets[xs] - ys;

  // 2b. products of the tile's nonzeros [ys, ye): every column id and value
  // load first, then the x gathers, then the LDS stores.
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
      if (FUSED) s_part[0] = val;
      else ws.head_val[tile] = val;
    } else {
      y[xs + first_row] = val;
    }
  }

  // 5. the row leaving the tile open
  const int xe_start = tile_rows > 0 ? s_rowend[tile_rows - 1] : s_coord[4];
  const bool open_right = xe < rows && xe_start < tile_nnz;
  if (tid == TPB - 1) {
    const float tot = f ? v : pv + v;
    if (FUSED) {
      s_part[1] = tot;
    } else {
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
  if (!FUSED) return;

  // In-kernel fix-up: wave 0 takes the open-left row, the last wave the
  // open-right row.  Each bumps the row's counter; the last participant sums
  // the slots of all participants in unit order.
  __syncthreads();
  const int part_wave_left = 0, part_wave_right = NW - 1;
  for (int which = 0; which < 2; ++which) {
    const bool mine = which == 0 ? (wave == part_wave_left && open_left)
                                 : (wave == part_wave_right && open_right);
    if (!mine) continue;
    const long long r = which == 0 ? xs : xe;
    const long long fr = which == 0 ? (long long)s_coord[4] + ys + xs
                                    : (long long)xe_start + ys + xe;
    const long long fr1 = which == 0 ? (long long)s_rowend[0] + ys + xs + 1
                                     : (long long)s_rowend[tile_rows] + ys + xe + 1;
    const long long t0 = fr / TI;
    const long long t1 = (fr1 - 1) / TI;
    const int n = (int)(t1 - t0 + 1);
    int last = 0;
    if (lane == 0) {
      // Publish this tile's partial, then count in (release + acquire).
      if (which == 0) ws.head_val[tile] = s_part[0];
      else ws.carry_val[tile] = s_part[1];
      int old = __hip_atomic_fetch_add(&ws.counter[t0], 1, __ATOMIC_ACQ_REL,
                                       __HIP_MEMORY_SCOPE_AGENT);
      last = old == n - 1;
    }
    last = __shfl(last, 0, 64);
    if (last) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      const float s = segment_is_short(t0, t1)
                          ? short_segment_sum(ws.carry_val, ws.head_val, t0, t1)
                          : wave_segment_sum(ws.carry_val, ws.head_val, t0, t1);
      if (lane == 0) {
        y[r] = s;
        __hip_atomic_store(&ws.counter[t0], 0, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

// Second-kernel fix-up (see the file comment).  Work-item t acts for the
// unit where an open row starts (seg_end[t] >= 0, written by the main
// kernel): a two-unit row is summed by the work-item itself, longer rows by
// the whole wavefront with the same fixed order as the in-kernel path.
__global__ __launch_bounds__(256) void spmv_fixup_kernel(long long num_units,
                                                         fixup_ws ws,
                                                         float* __restrict__ y) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  long long t1 = -1;
  int r = -1;
  if (t < num_units) {
    t1 = ws.seg_end[t];
    if (t1 >= 0) r = ws.carry_row[t];
  }
  if (t1 >= 0 && segment_is_short(t, t1))
    y[r] = short_segment_sum(ws.carry_val, ws.head_val, t, t1);
  unsigned long long todo = __ballot(t1 >= 0 && !segment_is_short(t, t1));
  while (todo) {
    const int src = __ffsll((long long)todo) - 1;
    todo &= todo - 1;
    const long long a = __shfl(t, src, 64);
    const long long b = __shfl(t1, src, 64);
    const int rr = __shfl(r, src, 64);
    const float v = wave_segment_sum(ws.carry_val, ws.head_val, a, b);
    if (lane == src) y[rr] = v;
  }
}

// Merge-path partition by scatter: work-item r writes the coordinate of
// every tile boundary t * items_per_tile that falls in row r's diagonals
// [f(r), f(r+1)), which is (r, t * items_per_tile - r); one extra work-item
// writes the end of the path (rows, nnz) at t = num_coords - 1.  Reads the
// offsets once, in order; equal to the search at every boundary because
// MP(d).x = max{k : f(k) <= d}.
__global__ __launch_bounds__(256) void mp_partition_kernel(
    const int* __restrict__ offsets, int rows, int nnz, long long items_per_tile,
    long long num_coords, loops_coord_t* __restrict__ coords) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > rows) return;
  if (r == rows) {
    coords[num_coords - 1] = loops_coord_t{(unsigned)rows, (unsigned)nnz};
    return;
  }
  const long long a = (long long)offsets[r] + r;
  const long long b = (long long)offsets[r + 1] + r + 1;
  for (long long t = (a + items_per_tile - 1) / items_per_tile;
       t * items_per_tile < b && t < num_coords - 1; ++t)
    coords[t] = loops_coord_t{(unsigned)r, (unsigned)(t * items_per_tile - r)};
}

// work_oriented partition (work_oriented.hxx:150-167) for num_threads
// threads: start and end coordinate of each.
__global__ void wo_partition_kernel(const int* __restrict__ offsets, int rows,
                                    int nnz, long long num_threads,
                                    loops_coord_t* __restrict__ coords) {
  long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= num_threads) return;
  long long total = (long long)rows + nnz;
  long long wpt = (total + num_threads - 1) / num_threads;
  long long upper = wpt * tid < total ? wpt * tid : total;
  long long lower = upper + wpt < total ? upper + wpt : total;
  auto st = loops::search::_binary_search<long long>(upper, offsets + 1, 0LL,
                                                     (long long)rows, (long long)nnz);
  auto en = loops::search::_binary_search<long long>(lower, offsets + 1, 0LL,
                                                     (long long)rows, (long long)nnz);
  coords[2 * tid] = loops_coord_t{st.x, st.y};
  coords[2 * tid + 1] = loops_coord_t{en.x, en.y};
}

// ---------------------------------------------------------------------------
// work_oriented SpMV, written against loops::schedule::setup<work_oriented>.
// Rows closed inside a thread's share are stored directly, except a row the
// share entered already open; open ends go through the fix-up slots instead
// of the reference's store/atomicAdd pair (which can race across blocks).
// ---------------------------------------------------------------------------
template <int TPB, typename V>
__global__ __launch_bounds__(TPB) void spmv_work_oriented_kernel(
    const int* __restrict__ offsets, const int* __restrict__ indices,
    const V* __restrict__ values, const float* __restrict__ x,
    float* __restrict__ y, int rows, int nnz, int* __restrict__ carry_row,
    int* __restrict__ seg_end, float* __restrict__ carry_val,
    float* __restrict__ head_val) {
  using setup_t = loops::schedule::setup<algorithms_t::work_oriented, TPB, 1,
                                         int, int, long long, long long>;
  setup_t config(const_cast<int*>(offsets), rows, nnz);
  auto map = config.init();
  const int sx = map.first.first, sy = map.first.second;
  const int ex = map.second.first, ey = map.second.second;
  const bool open_left = sx < ex && offsets[sx] < sy;

  float head = 0.f;
  for (auto row : config.tiles(map)) {
    float sum = 0.f;
    for (auto nz : config.atoms(row, map))
      sum += value_traits<V>::load1(values + nz) * x[ld_idx(indices + nz)];
    if (row == sx && open_left) head = sum;
    else y[row] = sum;
  }
  float rem = 0.f;
  for (auto row : config.remainder_tiles(map)) {
    (void)row;
    for (auto nz : config.remainder_atoms(map))
      rem += value_traits<V>::load1(values + nz) * x[ld_idx(indices + nz)];
  }
  long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool open_right = ex < rows && offsets[ex] < ey;
  carry_row[u] = open_right ? ex : -1;
  carry_val[u] = rem;
  head_val[u] = head;
  int end_unit = -1;
  if (open_right) {
    const long long wpt = config.work_per_thread_count();
    const long long fr = (long long)offsets[ex] + ex;
    if (fr / wpt == u) end_unit = (int)(((long long)offsets[ex + 1] + ex) / wpt);
  }
  seg_end[u] = end_unit;
}

// ---------------------------------------------------------------------------
// group_mapped SpMV, written against loops::schedule::setup<group_mapped>.
// A group of TPT lanes owns TPT rows; its nonzeros are visited TPT at a time
// in CSR order.  Instead of an atomicAdd per nonzero, each chunk is reduced
// by row with a segmented scan and the run totals are added, chunk by chunk,
// to per-row LDS accumulators, then stored once.
// ---------------------------------------------------------------------------
template <int TPB, int TPT, typename V>
__global__ __launch_bounds__(TPB) void spmv_group_mapped_kernel(
    const int* __restrict__ offsets, const int* __restrict__ indices,
    const V* __restrict__ values, const float* __restrict__ x,
    float* __restrict__ y, int rows, int nnz) {
  using setup_t = loops::schedule::setup<algorithms_t::group_mapped, TPB, TPT,
                                         int, int, long long, long long>;
  constexpr int NW = TPB / 64;
  __shared__ typename setup_t::storage_t storage;
  __shared__ float s_acc[TPB];
  __shared__ int s_key[TPB];
  __shared__ float s_wv[NW > 0 ? NW : 1];
  __shared__ int s_wf[NW > 0 ? NW : 1];

  setup_t config(storage, const_cast<int*>(offsets), rows, nnz);
  auto p = config.partition();
  auto atoms = config.atom_accessor(p);
  const int aggregate = *atoms.end();
  const int rank = (int)p.thread_rank();
  const int gbase = (int)p.meta_group_rank() * TPT;
  s_acc[threadIdx.x] = 0.f;
  p.sync();

  for (int base = 0; base < aggregate; base += TPT) {
    int v_atom = base + rank;
    int key = TPT;  // sentinel for lanes past the group's nonzeros
    float prod = 0.f;
    if (v_atom < aggregate) {
      int v_tile = config.tile_accessor(v_atom, p);
      int row = config.tile_id(v_tile, p);
      int nz = config.atom_id(v_atom, row, v_tile, p);
      prod = value_traits<V>::load1(values + nz) * x[ld_idx(indices + nz)];
      key = v_tile;
    }
    float v = prod;
    int head, tail;
    if constexpr (TPT <= 64) {
      int kp = __shfl_up(key, 1, TPT);
      int kn = __shfl_down(key, 1, TPT);
      head = (rank == 0 || kp != key) ? 1 : 0;
      tail = (rank == TPT - 1 || kn != key) ? 1 : 0;
      int f = head;
      group_seg_scan<TPT>(v, f);
    } else {
      s_key[threadIdx.x] = key;
      __syncthreads();
      head = (threadIdx.x == 0 || s_key[threadIdx.x - 1] != key) ? 1 : 0;
      tail = (threadIdx.x == TPB - 1 || s_key[threadIdx.x + 1] != key) ? 1 : 0;
      int f = head;
      float pv;
      int pf;
      block_seg_scan<TPB>(v, f, s_wv, s_wf, pv, pf);
    }
    if (tail && key < TPT) s_acc[gbase + key] = s_acc[gbase + key] + v;
    p.sync();
  }
  if (rank < config.get_length(p)) {
    int row = storage.tiles_indices[gbase + rank];
    y[row] = s_acc[gbase + rank];
  }
}

// ---------------------------------------------------------------------------
// thread_mapped SpMV through loops::schedule::setup<thread_mapped> (one row
// per work-item, grid-stride) and the schedule-free original.  Both sum each
// row sequentially in CSR order, exactly like the CPU validator.
// ---------------------------------------------------------------------------
template <typename V>
__global__ void spmv_thread_mapped_kernel(
    loops::schedule::setup<algorithms_t::thread_mapped, 1, 1, int, int,
                           long long, long long> config,
    const int* __restrict__ indices, const V* __restrict__ values,
    const float* __restrict__ x, float* __restrict__ y) {
  for (auto row : config.tiles()) {
    float sum = 0.f;
    for (auto nz : config.atoms(row))
      sum += value_traits<V>::load1(values + nz) * x[ld_idx(indices + nz)];
    y[row] = sum;
  }
}

template <typename V>
__global__ void spmv_original_kernel(const int* __restrict__ offsets,
                                     const int* __restrict__ indices,
                                     const V* __restrict__ values,
                                     const float* __restrict__ x,
                                     float* __restrict__ y, int rows) {
  for (long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
       row < rows; row += (long long)gridDim.x * blockDim.x) {
    float sum = 0.f;
    for (int nz = offsets[row]; nz < offsets[row + 1]; ++nz)
      sum += value_traits<V>::load1(values + nz) * x[ld_idx(indices + nz)];
    y[row] = sum;
  }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
namespace {

inline hipError_t record(hipEvent_t e, hipStream_t s) {
  return e ? hipEventRecord(e, s) : hipSuccess;
}

template <int TPB, int IPT, typename V, bool FUSED>
hipError_t mp_launch_t(const csr_args& A, const float* x, float* y,
                       const loops_coord_t* coords, const fixup_ws& ws,
                       hipEvent_t ev0, hipEvent_t ev1, hipStream_t s) {
  const long long total = (long long)A.rows + A.nnz;
  const long long TI = (long long)TPB * IPT;
  const long long tiles = (total + TI - 1) / TI;
  hipError_t e = record(ev0, s);
  if (e != hipSuccess) return e;
  if (tiles > 0) {
    const V* vals = static_cast<const V*>(A.values);
    if (A.aligned)
      spmv_merge_path_kernel<TPB, IPT, V, true, FUSED><<<dim3((unsigned)tiles), dim3(TPB), 0, s>>>(
          A.offsets, A.indices, vals, x, y, A.rows, A.nnz, coords, ws);
    else
      spmv_merge_path_kernel<TPB, IPT, V, false, FUSED><<<dim3((unsigned)tiles), dim3(TPB), 0, s>>>(
          A.offsets, A.indices, vals, x, y, A.rows, A.nnz, coords, ws);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  e = record(ev1, s);
  if (e != hipSuccess) return e;
  if (!FUSED && tiles > 1) return launch_fixup(A.offsets, A.rows, TI, tiles, ws, y, s);
  return hipSuccess;
}

template <int TPB, int IPT>
hipError_t mp_dispatch_v(const csr_args& A, const float* x, float* y,
                         const loops_coord_t* coords, const fixup_ws& ws,
                         bool fused, hipEvent_t ev0, hipEvent_t ev1,
                         hipStream_t s) {
  switch (A.vtype) {
    case LOOPS_VALUE_F32:
      return fused ? mp_launch_t<TPB, IPT, float, true>(A, x, y, coords, ws, ev0, ev1, s)
                   : mp_launch_t<TPB, IPT, float, false>(A, x, y, coords, ws, ev0, ev1, s);
    case LOOPS_VALUE_BF16:
      return fused ? mp_launch_t<TPB, IPT, bf16_t, true>(A, x, y, coords, ws, ev0, ev1, s)
                   : mp_launch_t<TPB, IPT, bf16_t, false>(A, x, y, coords, ws, ev0, ev1, s);
    case LOOPS_VALUE_F16:
      return fused ? mp_launch_t<TPB, IPT, f16_t, true>(A, x, y, coords, ws, ev0, ev1, s)
                   : mp_launch_t<TPB, IPT, f16_t, false>(A, x, y, coords, ws, ev0, ev1, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

// Tunings compiled for merge_path_flat: (threads per block, items per thread).
#define LOOPS_MP_TUNINGS(X) \
  X(64, 3)                  \
  X(64, 8)                  \
  X(128, 4)                 \
  X(128, 5)                 \
  X(128, 8)                 \
  X(128, 12)                \
  X(128, 16)                \
  X(256, 2)                 \
  X(256, 3)                 \
  X(256, 4)                 \
  X(256, 5)                 \
  X(256, 6)                 \
  X(256, 8)                 \
  X(256, 10)                \
  X(256, 12)                \
  X(256, 16)                \
  X(512, 2)                 \
  X(512, 4)                 \
  X(512, 6)                 \
  X(512, 8)

bool mp_spmv_supported(int tpb, int ipt) {
#define X(T, I) \
  if (tpb == T && ipt == I) return true;
  LOOPS_MP_TUNINGS(X)
#undef X
  return false;
}

hipError_t launch_mp_spmv(const csr_args& A, const float* x, float* y, int tpb,
                          int ipt, const loops_coord_t* coords,
                          const fixup_ws& ws, bool fused, hipEvent_t ev0,
                          hipEvent_t ev1, hipStream_t s) {
#define X(T, I)          \
  if (tpb == T && ipt == I) \
    return mp_dispatch_v<T, I>(A, x, y, coords, ws, fused, ev0, ev1, s);
  LOOPS_MP_TUNINGS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_fixup(const int* offsets, int rows, long long unit_items,
                        long long num_units, const fixup_ws& ws, float* y,
                        hipStream_t s) {
  if (num_units <= 0) return hipSuccess;
  const int tpb = 256;
  long long grid = (num_units + tpb - 1) / tpb;
  (void)offsets;
  (void)rows;
  (void)unit_items;
  spmv_fixup_kernel<<<dim3((unsigned)grid), dim3(tpb), 0, s>>>(num_units, ws, y);
  return hipGetLastError();
}

hipError_t launch_mp_partition(const int* offsets, int rows, int nnz,
                               long long items_per_tile, long long num_coords,
                               loops_coord_t* coords, hipStream_t s) {
  if (num_coords <= 0) return hipSuccess;
  long long grid = ((long long)rows + 1 + 255) / 256;  // one work-item per row, +1
  mp_partition_kernel<<<dim3((unsigned)grid), dim3(256), 0, s>>>(
      offsets, rows, nnz, items_per_tile, num_coords, coords);
  return hipGetLastError();
}

hipError_t launch_wo_partition(const int* offsets, int rows, int nnz,
                               long long num_threads, loops_coord_t* coords,
                               hipStream_t s) {
  if (num_threads <= 0) return hipSuccess;
  long long grid = (num_threads + 255) / 256;
  wo_partition_kernel<<<dim3((unsigned)grid), dim3(256), 0, s>>>(
      offsets, rows, nnz, num_threads, coords);
  return hipGetLastError();
}

namespace {
template <int TPB, typename V>
hipError_t wo_launch_t(const csr_args& A, const float* x, float* y, int grid,
                       const fixup_ws& ws, hipEvent_t ev0, hipEvent_t ev1,
                       hipStream_t s) {
  hipError_t e = record(ev0, s);
  if (e != hipSuccess) return e;
  spmv_work_oriented_kernel<TPB, V><<<dim3(grid), dim3(TPB), 0, s>>>(
      A.offsets, A.indices, static_cast<const V*>(A.values), x, y, A.rows,
      A.nnz, ws.carry_row, ws.seg_end, ws.carry_val, ws.head_val);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = record(ev1, s);
  if (e != hipSuccess) return e;
  const long long units = (long long)grid * TPB;
  const long long total = (long long)A.rows + A.nnz;
  const long long wpt = (total + units - 1) / units;
  return launch_fixup(A.offsets, A.rows, wpt, units, ws, y, s);
}
template <int TPB>
hipError_t wo_dispatch_v(const csr_args& A, const float* x, float* y, int grid,
                         const fixup_ws& ws, hipEvent_t ev0, hipEvent_t ev1,
                         hipStream_t s) {
  switch (A.vtype) {
    case LOOPS_VALUE_F32: return wo_launch_t<TPB, float>(A, x, y, grid, ws, ev0, ev1, s);
    case LOOPS_VALUE_BF16: return wo_launch_t<TPB, bf16_t>(A, x, y, grid, ws, ev0, ev1, s);
    case LOOPS_VALUE_F16: return wo_launch_t<TPB, f16_t>(A, x, y, grid, ws, ev0, ev1, s);
  }
  return hipErrorInvalidValue;
}
}  // namespace

bool wo_spmv_supported(int tpb) { return tpb == 64 || tpb == 128 || tpb == 256; }

hipError_t launch_wo_spmv(const csr_args& A, const float* x, float* y, int tpb,
                          int grid, const fixup_ws& ws, hipEvent_t ev0,
                          hipEvent_t ev1, hipStream_t s) {
  if ((long long)A.rows + A.nnz == 0) {
    hipError_t e = record(ev0, s);
    return e != hipSuccess ? e : record(ev1, s);
  }
  switch (tpb) {
    case 64: return wo_dispatch_v<64>(A, x, y, grid, ws, ev0, ev1, s);
    case 128: return wo_dispatch_v<128>(A, x, y, grid, ws, ev0, ev1, s);
    case 256: return wo_dispatch_v<256>(A, x, y, grid, ws, ev0, ev1, s);
  }
  return hipErrorInvalidValue;
}

namespace {
template <int TPB, int TPT, typename V>
hipError_t gm_launch_t(const csr_args& A, const float* x, float* y,
                       hipEvent_t ev0, hipEvent_t ev1, hipStream_t s) {
  hipError_t e = record(ev0, s);
  if (e != hipSuccess) return e;
  long long grid = ((long long)A.rows + TPB - 1) / TPB;
  if (grid > 0) {
    spmv_group_mapped_kernel<TPB, TPT, V><<<dim3((unsigned)grid), dim3(TPB), 0, s>>>(
        A.offsets, A.indices, static_cast<const V*>(A.values), x, y, A.rows, A.nnz);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return record(ev1, s);
}
template <int TPB, int TPT>
hipError_t gm_dispatch_v(const csr_args& A, const float* x, float* y,
                         hipEvent_t ev0, hipEvent_t ev1, hipStream_t s) {
  switch (A.vtype) {
    case LOOPS_VALUE_F32: return gm_launch_t<TPB, TPT, float>(A, x, y, ev0, ev1, s);
    case LOOPS_VALUE_BF16: return gm_launch_t<TPB, TPT, bf16_t>(A, x, y, ev0, ev1, s);
    case LOOPS_VALUE_F16: return gm_launch_t<TPB, TPT, f16_t>(A, x, y, ev0, ev1, s);
  }
  return hipErrorInvalidValue;
}
}  // namespace

#define LOOPS_GM_TUNINGS(X) \
  X(64, 64)                 \
  X(128, 128)               \
  X(256, 256)               \
  X(128, 64)                \
  X(256, 64)                \
  X(256, 32)                \
  X(256, 16)                \
  X(256, 8)                 \
  X(256, 4)

bool gm_spmv_supported(int tpb, int tpt) {
#define X(T, G) \
  if (tpb == T && tpt == G) return true;
  LOOPS_GM_TUNINGS(X)
#undef X
  return false;
}

hipError_t launch_gm_spmv(const csr_args& A, const float* x, float* y, int tpb,
                          int tpt, hipEvent_t ev0, hipEvent_t ev1,
                          hipStream_t s) {
#define X(T, G)             \
  if (tpb == T && tpt == G) \
    return gm_dispatch_v<T, G>(A, x, y, ev0, ev1, s);
  LOOPS_GM_TUNINGS(X)
#undef X
  return hipErrorInvalidValue;
}

namespace {
template <typename V>
hipError_t tm_launch_t(const csr_args& A, const float* x, float* y, int tpb,
                       bool original, hipEvent_t ev0, hipEvent_t ev1,
                       hipStream_t s) {
  hipError_t e = record(ev0, s);
  if (e != hipSuccess) return e;
  long long grid = ((long long)A.rows + tpb - 1) / tpb;
  if (grid > 0) {
    const V* vals = static_cast<const V*>(A.values);
    if (original) {
      spmv_original_kernel<V><<<dim3((unsigned)grid), dim3(tpb), 0, s>>>(
          A.offsets, A.indices, vals, x, y, A.rows);
    } else {
      loops::schedule::setup<algorithms_t::thread_mapped, 1, 1, int, int,
                             long long, long long>
          config(const_cast<int*>(A.offsets), A.rows, A.nnz);
      spmv_thread_mapped_kernel<V><<<dim3((unsigned)grid), dim3(tpb), 0, s>>>(
          config, A.indices, vals, x, y);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return record(ev1, s);
}
}  // namespace

hipError_t launch_tm_spmv(const csr_args& A, const float* x, float* y, int tpb,
                          bool original, hipEvent_t ev0, hipEvent_t ev1,
                          hipStream_t s) {
  switch (A.vtype) {
    case LOOPS_VALUE_F32: return tm_launch_t<float>(A, x, y, tpb, original, ev0, ev1, s);
    case LOOPS_VALUE_BF16: return tm_launch_t<bf16_t>(A, x, y, tpb, original, ev0, ev1, s);
    case LOOPS_VALUE_F16: return tm_launch_t<f16_t>(A, x, y, tpb, original, ev0, ev1, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace loops_mi355x
