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
// additions on every run.  The fix-up is a second, small kernel.  (Doing it
// inside the merge-path kernel with last-arriver counters needs an
// agent-scope release and acquire per tile, i.e. an L2 write-back and
// invalidate on gfx950, which measured 4-5x slower overall.)
#include <hip/hip_runtime.h>

#include <loops/schedule.hxx>
#include <loops/util/search.hxx>

#include "device_common.hpp"
#include "loops_internal.hpp"
#include "merge_tile.hpp"

namespace loops_mi355x {

using loops::schedule::algorithms_t;

fixup_ws carve_fixup_ws(void* ws, long long units) {
  char* p = static_cast<char*>(ws);
  fixup_ws w;
  w.carry_row = reinterpret_cast<int*>(p);
  w.seg_end = reinterpret_cast<int*>(p + units * sizeof(int));
  w.carry_val = reinterpret_cast<float*>(p + units * 2 * sizeof(int));
  w.head_val = reinterpret_cast<float*>(p + units * (2 * sizeof(int) + sizeof(float)));
  return w;
}

template <int TPB, int IPT, typename V, bool VEC>
__global__ __launch_bounds__(TPB) void spmv_merge_path_kernel(
    const int* __restrict__ offsets, const int* __restrict__ indices,
    const V* __restrict__ values, const float* __restrict__ x,
    float* __restrict__ y, int rows, int nnz,
    const loops_coord_t* __restrict__ coords, fixup_ws ws) {
  __shared__ mp_smem<TPB, IPT> sm;
  mp_tile<TPB, IPT, V, VEC>(offsets, indices, values, x, y, rows, nnz,
                            (int)blockIdx.x, coords, ws, sm);
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

  // The (row, atom) sequence of config.tiles() / atoms() / remainder_atoms():
  // the share's nonzeros sy .. ey - 1 in order, row t in [sx, ex) closing at
  // offsets[t + 1], the rest belonging to row ex.  It is walked here as one
  // run of nonzeros, kWoBatch at a time: a batch's index, value and x loads
  // are all issued before its products are added, so a lane waits on memory
  // once per batch instead of once per nonzero (the reference's launch gives
  // each thread hundreds of nonzeros and the GPU only 4 wavefronts per CU).
  // The products and sums are the same operations in the same order.
  constexpr int kWoBatch = 16;
  float head = 0.f, sum = 0.f;
  int t = sx;
  long long t_end = t < ex ? (long long)offsets[t + 1] : (long long)ey;
  auto close_rows = [&](long long upto) {  // rows of [sx, ex) ending at or before upto
    while (t < ex && t_end <= upto) {
      if (t == sx && open_left) head = sum;
      else y[t] = sum;
      sum = 0.f;
      ++t;
      t_end = t < ex ? (long long)offsets[t + 1] : (long long)ey;
    }
  };
  for (long long nz = sy; nz < ey; nz += kWoBatch) {
    const int n = ey - nz < kWoBatch ? (int)(ey - nz) : kWoBatch;
    int col[kWoBatch];
    float val[kWoBatch], xv[kWoBatch];
#pragma unroll
    for (int j = 0; j < kWoBatch; ++j) {
      col[j] = j < n ? ld_idx(indices + nz + j) : 0;
      val[j] = j < n ? value_traits<V>::load1(values + nz + j) : 0.f;
    }
#pragma unroll
    for (int j = 0; j < kWoBatch; ++j) xv[j] = j < n ? x[col[j]] : 0.f;
#pragma unroll
    for (int j = 0; j < kWoBatch; ++j) {
      if (j < n) {
        close_rows(nz + j);
        sum += val[j] * xv[j];
      }
    }
  }
  close_rows(ey);
  const float rem = sum;
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

  // Groups of several wavefronts scan across the whole block (workgroup
  // barriers inside), so every group runs the block's largest trip count;
  // keys are then block-wide (group base + group-local row) and a lane past
  // its group's nonzeros takes a sentinel of its own group.
  int trip = aggregate;
  if constexpr (TPT > 64 && TPT < TPB) {
    for (int g = 0; g < TPB / TPT; ++g)
      trip = storage.tile_aggregates[g] > trip ? storage.tile_aggregates[g] : trip;
  }
  constexpr int kSentinel = TPT > 64 ? TPB : TPT;
  // kGmAhead chunks at a time: their atoms are mapped and all their loads
  // issued (branch-free: a lane past its group's nonzeros loads nonzero 0
  // and drops the product) before the first is reduced, so a group waits on
  // memory once per kGmAhead chunks; the chunks are then reduced and added
  // in order, as one at a time would be.
  constexpr int kGmAhead = 8;
  for (int base0 = 0; base0 < trip; base0 += kGmAhead * TPT) {
    int keys[kGmAhead], nzs[kGmAhead];
    float prods[kGmAhead];
#pragma unroll
    for (int j = 0; j < kGmAhead; ++j) {
      int v_atom = base0 + j * TPT + rank;  // the accessors take references
      keys[j] = kSentinel + (TPT > 64 ? (int)p.meta_group_rank() : 0);
      nzs[j] = 0;
      if (v_atom < aggregate) {
        int v_tile = config.tile_accessor(v_atom, p);
        int row = config.tile_id(v_tile, p);
        nzs[j] = config.atom_id(v_atom, row, v_tile, p);
        keys[j] = TPT > 64 ? gbase + v_tile : v_tile;
      }
    }
    int cols[kGmAhead];
    float vals[kGmAhead];
#pragma unroll
    for (int j = 0; j < kGmAhead; ++j) {
      cols[j] = ld_idx(indices + nzs[j]);
      vals[j] = value_traits<V>::load1(values + nzs[j]);
    }
#pragma unroll
    for (int j = 0; j < kGmAhead; ++j) {
      const float xv = x[cols[j]];
      prods[j] = base0 + j * TPT + rank < aggregate ? vals[j] * xv : 0.f;
    }
#pragma unroll
    for (int j = 0; j < kGmAhead; ++j) {
    if (base0 + j * TPT >= trip) break;
    const int key = keys[j];
    float v = prods[j];
    int head, tail;
    if constexpr (TPT <= 64) {
      int kp = __shfl_up(key, 1, TPT);
      int kn = __shfl_down(key, 1, TPT);
      head = (rank == 0 || kp != key) ? 1 : 0;
      tail = (rank == TPT - 1 || kn != key) ? 1 : 0;
      int f = head;
      group_seg_scan<TPT>(v, f);
      if (tail && key < TPT) s_acc[gbase + key] = s_acc[gbase + key] + v;
    } else {
      s_key[threadIdx.x] = key;
      __syncthreads();
      head = (threadIdx.x == 0 || s_key[threadIdx.x - 1] != key) ? 1 : 0;
      tail = (threadIdx.x == TPB - 1 || s_key[threadIdx.x + 1] != key) ? 1 : 0;
      int f = head;
      float pv;
      int pf;
      block_seg_scan<TPB>(v, f, s_wv, s_wf, pv, pf);
      if (tail && key < TPB) s_acc[key] = s_acc[key] + v;
    }
    p.sync();
    }
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
// Sum over nonzeros b .. e - 1 of value * x[column], added front to back
// (the order of the reference's CPU validator).  The loads of kSeqBatch
// nonzeros are all issued before their products are added, so a lane waits
// on memory once per batch; the operations and their order are unchanged.
constexpr int kSeqBatch = 16;
template <typename V>
__device__ __forceinline__ float seq_row_dot(const int* __restrict__ indices,
                                             const V* __restrict__ values,
                                             const float* __restrict__ x, long long b,
                                             long long e) {
  float sum = 0.f;
  for (long long nz = b; nz < e; nz += kSeqBatch) {
    const int n = e - nz < kSeqBatch ? (int)(e - nz) : kSeqBatch;
    int col[kSeqBatch];
    float val[kSeqBatch], xv[kSeqBatch];
#pragma unroll
    for (int j = 0; j < kSeqBatch; ++j) {
      col[j] = j < n ? ld_idx(indices + nz + j) : 0;
      val[j] = j < n ? value_traits<V>::load1(values + nz + j) : 0.f;
    }
#pragma unroll
    for (int j = 0; j < kSeqBatch; ++j) xv[j] = j < n ? x[col[j]] : 0.f;
#pragma unroll
    for (int j = 0; j < kSeqBatch; ++j)
      if (j < n) sum += val[j] * xv[j];
  }
  return sum;
}

template <typename V>
__global__ void spmv_thread_mapped_kernel(
    loops::schedule::setup<algorithms_t::thread_mapped, 1, 1, int, int,
                           long long, long long> config,
    const int* __restrict__ indices, const V* __restrict__ values,
    const float* __restrict__ x, float* __restrict__ y) {
  for (auto row : config.tiles()) {
    const auto atoms = config.atoms(row);  // the row's nonzeros, in CSR order
    y[row] = seq_row_dot(indices, values, x, *atoms.begin(), *atoms.end());
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
    y[row] = seq_row_dot(indices, values, x, offsets[row], offsets[row + 1]);
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
          A.offsets, A.indices, vals, x, y, A.rows, A.nnz, coords, ws);
    else
      spmv_merge_path_kernel<TPB, IPT, V, false><<<dim3((unsigned)tiles), dim3(TPB), 0, s>>>(
          A.offsets, A.indices, vals, x, y, A.rows, A.nnz, coords, ws);
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
    default:
      break;
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
    default: break;
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
    default: break;
  }
  return hipErrorInvalidValue;
}
}  // namespace

#define LOOPS_GM_TUNINGS(X) \
  X(64, 64)                 \
  X(128, 128)               \
  X(256, 256)               \
  X(256, 128)               \
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
    default: break;
  }
  return hipErrorInvalidValue;
}

}  // namespace loops_mi355x
