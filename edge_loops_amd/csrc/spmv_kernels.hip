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
// launch_fixup: each unit records the row it leaves open (carry_row), its
// partial sum of that row (carry_val) and its partial sum of the row it
// entered already open (head_val).  The unit in which a row starts sums the
// unit partials of that row in unit order, so every y[row] is produced by
// one fixed sequence of fp32 additions on every run.
#include <hip/hip_runtime.h>

#include <loops/schedule.hxx>
#include <loops/util/search.hxx>

#include "device_common.hpp"
#include "loops_internal.hpp"

namespace loops_mi355x {

using loops::schedule::algorithms_t;

fixup_ws carve_fixup_ws(void* ws, long long units) {
  char* p = static_cast<char*>(ws);
  fixup_ws w;
  w.carry_row = reinterpret_cast<int*>(p);
  w.carry_val = reinterpret_cast<float*>(p + units * sizeof(int));
  w.head_val = reinterpret_cast<float*>(p + units * (sizeof(int) + sizeof(float)));
  return w;
}

// ---------------------------------------------------------------------------
// Merge-path SpMV.
//
// One workgroup per merge tile of TI = TPB * IPT steps (rows + nonzeros).
//  1. Tile end points: read from the partition, or found by two wavefronts,
//     each running a 64-way cooperative search (loops::search::wave_search).
//  2. The row ends of the tile's rows go to LDS; the tile's nonzeros are
//     streamed with 16-byte non-temporal loads (values and column ids) and
//     their products a_ij * x_j go to LDS, so every HBM byte of A is read
//     once, coalesced, and x keeps the L2.
//  3. Each lane finds its start with a search in LDS and walks IPT merge
//     steps (consume a product or close a row), writing closed rows to LDS.
//  4. A workgroup segmented scan carries each lane's trailing partial into
//     the next lane's first closed row.
//  5. Closed rows are stored coalesced; the open ends go to the fix-up slots.
// ---------------------------------------------------------------------------
template <int TPB, int IPT, typename V, bool VEC>
__global__ __launch_bounds__(TPB) void spmv_merge_path_kernel(
    const int* __restrict__ offsets, const int* __restrict__ indices,
    const V* __restrict__ values, const float* __restrict__ x,
    float* __restrict__ y, int rows, int nnz,
    const loops_coord_t* __restrict__ coords, int* __restrict__ carry_row,
    float* __restrict__ carry_val, float* __restrict__ head_val) {
  constexpr int TI = TPB * IPT;
  constexpr int NW = TPB / 64;
  __shared__ int s_rowend[TI + 1];
  __shared__ __align__(16) float s_prod[TI + 4];
  __shared__ float s_y[TI];
  __shared__ int s_coord[5];
  __shared__ float s_wv[NW];
  __shared__ int s_wf[NW];

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

  // 2b. products of the tile's nonzeros [ys, ye)
  const int ybase = ys & ~3;
  const int pofs = ys - ybase;
  if (VEC) {
    for (int j = ybase + 4 * tid; j < ye; j += 4 * TPB) {
      if (j >= ys && j + 4 <= ye) {
        i32x4 c = ld_idx4(indices + j);
        f32x4 v = value_traits<V>::load4(values + j);
        f32x4 p;
        p.x = v.x * x[c.x];
        p.y = v.y * x[c.y];
        p.z = v.z * x[c.z];
        p.w = v.w * x[c.w];
        *reinterpret_cast<f32x4*>(&s_prod[j - ybase]) = p;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          int jj = j + e;
          if (jj >= ys && jj < ye)
            s_prod[jj - ybase] =
                value_traits<V>::load1(values + jj) * x[ld_idx(indices + jj)];
        }
      }
    }
  } else {
    for (int j = ys + tid; j < ye; j += TPB)
      s_prod[j - ybase] = value_traits<V>::load1(values + j) * x[ld_idx(indices + j)];
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
  int first_emit = -1;
#pragma unroll
  for (int s = 0; s < IPT; ++s) {
    if (s < steps) {
      if (yl < s_rowend[xl]) {
        run += s_prod[yl + pofs];
        ++yl;
      } else {
        s_y[xl] = run;
        if (first_emit < 0) first_emit = xl;
        run = 0.f;
        ++xl;
      }
    }
  }

  // 4. carry each lane's trailing partial into the next closed row
  float v = run;
  int f = first_emit >= 0 ? 1 : 0;
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
  float excl = lane == 0 ? pv : (ef ? ev : pv + ev);
  if (first_emit >= 0) s_y[first_emit] = excl + s_y[first_emit];
  __syncthreads();

  // 5. store closed rows; record the open ends
  const bool open_left = tile_rows > 0 && s_coord[4] < 0;
  for (int i = tid; i < tile_rows; i += TPB) {
    if (i == 0 && open_left) continue;
    y[xs + i] = s_y[i];
  }
  if (tid == TPB - 1) {
    float tot = f ? v : pv + v;
    const int xe_start = tile_rows > 0 ? s_rowend[tile_rows - 1] : s_coord[4];
    carry_row[tile] = (xe < rows && xe_start < tile_nnz) ? xe : -1;
    carry_val[tile] = tot;
    head_val[tile] = open_left ? s_y[0] : 0.f;
  }
}

// Finishes rows that cross work-unit ends (see the file comment).  Unit u
// covers merge diagonals [u * unit_items, (u + 1) * unit_items).
__global__ void spmv_fixup_kernel(const int* __restrict__ offsets, int rows,
                                  long long unit_items, long long num_units,
                                  const int* __restrict__ carry_row,
                                  const float* __restrict__ carry_val,
                                  const float* __restrict__ head_val,
                                  float* __restrict__ y) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= num_units) return;
  int r = carry_row[t];
  if (r < 0 || r >= rows) return;
  long long fr = (long long)offsets[r] + r;
  if (fr / unit_items != t) return;  // not the unit where row r starts
  long long fr1 = (long long)offsets[r + 1] + r + 1;
  long long t_end = (fr1 - 1) / unit_items;
  float s = carry_val[t];
  for (long long u = t + 1; u < t_end; ++u) s += carry_val[u];
  s += head_val[t_end];
  y[r] = s;
}

// One wavefront per coordinate: coords[c] = MP(min(c * items_per_tile,
// rows + nnz)).
__global__ __launch_bounds__(256) void mp_partition_kernel(
    const int* __restrict__ offsets, int rows, int nnz, long long items_per_tile,
    long long num_coords, loops_coord_t* __restrict__ coords) {
  long long c = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (c >= num_coords) return;  // wave-uniform
  long long total = (long long)rows + nnz;
  long long d = c * items_per_tile;
  if (d > total) d = total;
  auto xy = loops::search::wave_search(offsets, rows, nnz, d);
  if ((threadIdx.x & 63) == 0) coords[c] = loops_coord_t{xy.x, xy.y};
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
    float* __restrict__ carry_val, float* __restrict__ head_val) {
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
  carry_row[u] = (ex < rows && offsets[ex] < ey) ? ex : -1;
  carry_val[u] = rem;
  head_val[u] = head;
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

template <int TPB, int IPT, typename V>
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
      spmv_merge_path_kernel<TPB, IPT, V, true><<<dim3((unsigned)tiles), dim3(TPB), 0, s>>>(
          A.offsets, A.indices, vals, x, y, A.rows, A.nnz, coords, ws.carry_row,
          ws.carry_val, ws.head_val);
    else
      spmv_merge_path_kernel<TPB, IPT, V, false><<<dim3((unsigned)tiles), dim3(TPB), 0, s>>>(
          A.offsets, A.indices, vals, x, y, A.rows, A.nnz, coords, ws.carry_row,
          ws.carry_val, ws.head_val);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  e = record(ev1, s);
  if (e != hipSuccess) return e;
  if (tiles > 1) return launch_fixup(A.offsets, A.rows, TI, tiles, ws, y, s);
  return hipSuccess;
}

template <int TPB, int IPT>
hipError_t mp_dispatch_v(const csr_args& A, const float* x, float* y,
                         const loops_coord_t* coords, const fixup_ws& ws,
                         hipEvent_t ev0, hipEvent_t ev1, hipStream_t s) {
  switch (A.vtype) {
    case LOOPS_VALUE_F32:
      return mp_launch_t<TPB, IPT, float>(A, x, y, coords, ws, ev0, ev1, s);
    case LOOPS_VALUE_BF16:
      return mp_launch_t<TPB, IPT, bf16_t>(A, x, y, coords, ws, ev0, ev1, s);
    case LOOPS_VALUE_F16:
      return mp_launch_t<TPB, IPT, f16_t>(A, x, y, coords, ws, ev0, ev1, s);
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
                          const fixup_ws& ws, hipEvent_t ev0, hipEvent_t ev1,
                          hipStream_t s) {
#define X(T, I)          \
  if (tpb == T && ipt == I) \
    return mp_dispatch_v<T, I>(A, x, y, coords, ws, ev0, ev1, s);
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
  spmv_fixup_kernel<<<dim3((unsigned)grid), dim3(tpb), 0, s>>>(
      offsets, rows, unit_items, num_units, ws.carry_row, ws.carry_val,
      ws.head_val, y);
  return hipGetLastError();
}

hipError_t launch_mp_partition(const int* offsets, int rows, int nnz,
                               long long items_per_tile, long long num_coords,
                               loops_coord_t* coords, hipStream_t s) {
  if (num_coords <= 0) return hipSuccess;
  long long grid = (num_coords + 3) / 4;  // 4 wavefronts per 256-lane block
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
      A.nnz, ws.carry_row, ws.carry_val, ws.head_val);
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
