// SpMV y = A x for the value types the reference's host entries are
// instantiated with besides float: double (benchmarks/spmv/work_oriented.cu
// :53, value_types = int, float, double) and int.  A's values, x and y all
// have the one type; every product and sum is in that type (int: 32-bit
// wrap-around, so any summation order gives the same bits).
//
// The five schedules keep the partitions of the fp32 kernels and of the
// reference (merge_path_flat: tiles of TPB x IPT steps, 64 x 3 for 8-byte
// values and 128 x 5 for 4-byte ones by default, merge_path_flat.cuh:
// 102-103; work_oriented: 2 x CU count blocks of 128; group_mapped:
// block_mapped<128>; thread_mapped / original: a row per work-item) and
// their fixed-order sums: rows split between work units are finished by a
// fix-up that adds the units' partial sums in unit order, so results are
// deterministic.  Loads are scalar and the tiles are not staged through LDS:
// this path is for type coverage, the fp32 kernels are the tuned ones.
#include <hip/hip_runtime.h>

#include <cstdint>

#include <loops/schedule.hxx>
#include <loops/util/search.hxx>

#include "device_common.hpp"
#include "loops_internal.hpp"

namespace loops_mi355x {

using loops::schedule::algorithms_t;

// Fix-up slots of the typed kernels: per work unit the open row, the unit
// holding its end (or -1), and the carried and head partial sums.
template <typename S>
struct typed_fixup {
  int* carry_row;
  int* seg_end;
  S* carry_val;
  S* head_val;
};

template <typename S>
typed_fixup<S> carve_typed_fixup(void* ws, long long units) {
  char* p = static_cast<char*>(ws);
  const size_t bi = (((size_t)units * sizeof(int)) + 255) & ~size_t(255);
  const size_t bs = (((size_t)units * sizeof(S)) + 255) & ~size_t(255);
  typed_fixup<S> w;
  w.carry_row = reinterpret_cast<int*>(p);
  w.seg_end = reinterpret_cast<int*>(p + bi);
  w.carry_val = reinterpret_cast<S*>(p + 2 * bi);
  w.head_val = reinterpret_cast<S*>(p + 2 * bi + bs);
  return w;
}

size_t typed_fixup_bytes(long long units, int value_bytes) {
  const size_t bi = (((size_t)units * sizeof(int)) + 255) & ~size_t(255);
  const size_t bs = (((size_t)units * value_bytes) + 255) & ~size_t(255);
  return 2 * bi + 2 * bs;
}

namespace {

// Segmented inclusive scan over groups of W lanes (W <= 64): a lane with f
// set starts a new segment.
template <int W, typename S>
__device__ __forceinline__ void typed_seg_scan(S& v, int& f) {
  const int lane = threadIdx.x % W;
#pragma unroll
  for (int d = 1; d < W; d <<= 1) {
    const S vu = __shfl_up(v, d, W);
    const int fu = __shfl_up(f, d, W);
    if (lane >= d) {
      if (!f) v = vu + v;
      f |= fu;
    }
  }
}

// The same over the workgroup: wavefront scans, then the totals of the
// wavefronts before a lane's own, in wavefront order.  pv / pf: the carry
// into the lane's wavefront.
template <int TPB, typename S>
__device__ __forceinline__ void typed_block_scan(S& v, int& f, S* s_v, int* s_f, S& pv,
                                                 int& pf) {
  constexpr int NW = TPB / 64;
  typed_seg_scan<64>(v, f);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 63) {
    s_v[wave] = v;
    s_f[wave] = f;
  }
  __syncthreads();
  pv = S(0);
  pf = 0;
  for (int w = 0; w < wave && w < NW; ++w) {
    pv = s_f[w] ? s_v[w] : pv + s_v[w];
    pf |= s_f[w];
  }
  if (!f) v = pv + v;
  f |= pf;
  __syncthreads();
}

// carry[t0] + ... + carry[t1 - 1] + head[t1]: lane l adds elements l, l+64,
// ... in sequence, the 64 lane sums are combined by a fixed butterfly.  A
// segment of one carry is added by its own work-item in the same order the
// butterfly gives it (exact zeros added).
template <typename S>
__device__ __forceinline__ S typed_segment_sum(const S* __restrict__ carry,
                                               const S* __restrict__ head, long long t0,
                                               long long t1) {
  const int lane = threadIdx.x & 63;
  S s = S(0);
  for (long long u = t0 + lane; u < t1; u += 64) s = s + carry[u];
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) s = s + __shfl_xor(s, m, 64);
  return s + head[t1];
}

template <typename S>
__global__ __launch_bounds__(256) void typed_fixup_kernel(long long num_units,
                                                          typed_fixup<S> ws,
                                                          S* __restrict__ y) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  long long t1 = -1;
  int r = -1;
  if (t < num_units) {
    t1 = ws.seg_end[t];
    if (t1 >= 0) r = ws.carry_row[t];
  }
  if (t1 >= 0 && t1 - t <= 1) y[r] = (t < t1 ? ws.carry_val[t] : S(0)) + ws.head_val[t1];
  unsigned long long todo = __ballot(t1 >= 0 && t1 - t > 1);
  while (todo) {
    const int src = __ffsll((long long)todo) - 1;
    todo &= todo - 1;
    const long long a = __shfl(t, src, 64);
    const long long b = __shfl(t1, src, 64);
    const int rr = __shfl(r, src, 64);
    const S v = typed_segment_sum(ws.carry_val, ws.head_val, a, b);
    if (lane == src) y[rr] = v;
  }
}

// merge_path_flat: one workgroup per merge tile of TPB * IPT steps.  The
// tile's end points (precomputed or searched), its row ends in LDS, each
// lane's start by a search in LDS (merge_path_flat.hxx:321-385), IPT steps
// of the consume rule (merge_path_flat.cuh:75-81), a workgroup segmented
// scan carrying partial sums between lanes, and the rows entering and
// leaving the tile open to the fix-up slots.
template <int TPB, int IPT, typename S>
__global__ __launch_bounds__(TPB) void typed_merge_path_kernel(
    const int* __restrict__ offsets, const int* __restrict__ indices,
    const S* __restrict__ values, const S* __restrict__ x, S* __restrict__ y, int rows,
    int nnz, const loops_coord_t* __restrict__ coords, typed_fixup<S> ws) {
  constexpr int TI = TPB * IPT;
  __shared__ int s_rowend[TI + 1];
  __shared__ int s_coord[5];
  __shared__ S s_v[TPB / 64];
  __shared__ int s_f[TPB / 64];
  const int tile = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long long total = (long long)rows + nnz;
  long long d0 = (long long)tile * TI;
  if (d0 > total) d0 = total;
  long long d1 = d0 + TI;
  if (d1 > total) d1 = total;
  if (coords != nullptr) {
    if (tid < 2) {
      const loops_coord_t c = coords[tile + tid];
      s_coord[2 * tid] = (int)c.x;
      s_coord[2 * tid + 1] = (int)c.y;
    }
  } else if (wave == 0) {
    const auto c0 = loops::search::wave_search(offsets, rows, nnz, d0);
    const auto c1 = loops::search::wave_search(offsets, rows, nnz, d1);
    if (lane == 0) {
      s_coord[0] = (int)c0.x;
      s_coord[1] = (int)c0.y;
      s_coord[2] = (int)c1.x;
      s_coord[3] = (int)c1.y;
    }
  }
  __syncthreads();
  const int xs = s_coord[0], ys = s_coord[1], xe = s_coord[2], ye = s_coord[3];
  const int tile_rows = xe - xs, tile_nnz = ye - ys;
  const int tile_items = (int)(d1 - d0);
  for (int i = tid; i <= tile_rows; i += TPB) {
    int r = xs + 1 + i;
    if (r > rows) r = rows;
    s_rowend[i] = offsets[r] - ys;
  }
  if (tid == 0) s_coord[4] = offsets[xs] - ys;
  __syncthreads();

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
      const int step = count >> 1, idx = first + step;
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
  S run = S(0), first_val = S(0);
  int first_row = -1;
  for (int st = 0; st < steps; ++st) {
    if (yl < s_rowend[xl]) {
      run = run + values[ys + yl] * x[indices[ys + yl]];
      ++yl;
    } else {
      if (first_row < 0) {
        first_row = xl;
        first_val = run;
      } else {
        y[xs + xl] = run;
      }
      run = S(0);
      ++xl;
    }
  }
  S v = run;
  int f = first_row >= 0 ? 1 : 0;
  S pv;
  int pf;
  typed_block_scan<TPB>(v, f, s_v, s_f, pv, pf);
  // exclusive value: the previous lane's inclusive one
  const S ev = __shfl_up(v, 1, 64);
  const int ef = __shfl_up(f, 1, 64);
  (void)ef;
  const S excl = lane == 0 ? pv : ev;
  const bool open_left = tile_rows > 0 && s_coord[4] < 0;
  if (first_row >= 0) {
    const S val = excl + first_val;
    if (first_row == 0 && open_left) ws.head_val[tile] = val;
    else y[xs + first_row] = val;
  }
  const int xe_start = tile_rows > 0 ? s_rowend[tile_rows - 1] : s_coord[4];
  const bool open_right = xe < rows && xe_start < tile_nnz;
  if (tid == TPB - 1) {
    ws.carry_row[tile] = open_right ? xe : -1;
    ws.carry_val[tile] = v;
    int seg_end = -1;
    if (open_right) {
      const long long fr = (long long)xe_start + ys + xe;
      if (fr / TI == tile) seg_end = (int)(((long long)s_rowend[tile_rows] + ys + xe) / TI);
    }
    ws.seg_end[tile] = seg_end;
  }
}

// work_oriented through setup<work_oriented> (work_oriented.hxx:150-195):
// rows closed in a share stored, except one the share entered open; the
// open ends go to the fix-up slots (unit = work-item).
template <int TPB, typename S>
__global__ __launch_bounds__(TPB) void typed_work_oriented_kernel(
    const int* __restrict__ offsets, const int* __restrict__ indices,
    const S* __restrict__ values, const S* __restrict__ x, S* __restrict__ y, int rows,
    int nnz, typed_fixup<S> ws) {
  using setup_t = loops::schedule::setup<algorithms_t::work_oriented, TPB, 1, int, int,
                                         long long, long long>;
  setup_t config(const_cast<int*>(offsets), rows, nnz);
  auto map = config.init();
  const int sx = map.first.first, sy = map.first.second;
  const int ex = map.second.first, ey = map.second.second;
  const bool open_left = sx < ex && offsets[sx] < sy;
  S head = S(0);
  for (auto row : config.tiles(map)) {
    S sum = S(0);
    for (auto nz : config.atoms(row, map)) sum = sum + values[nz] * x[indices[nz]];
    if (row == sx && open_left) head = sum;
    else y[row] = sum;
  }
  S rem = S(0);
  for (auto row : config.remainder_tiles(map)) {
    (void)row;
    for (auto nz : config.remainder_atoms(map)) rem = rem + values[nz] * x[indices[nz]];
  }
  const long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool open_right = ex < rows && offsets[ex] < ey;
  ws.carry_row[u] = open_right ? ex : -1;
  ws.carry_val[u] = rem;
  ws.head_val[u] = head;
  int end_unit = -1;
  if (open_right) {
    const long long wpt = config.work_per_thread_count();
    const long long fr = (long long)offsets[ex] + ex;
    if (fr / wpt == u) end_unit = (int)(((long long)offsets[ex + 1] + ex) / wpt);
  }
  ws.seg_end[u] = end_unit;
}

// group_mapped through setup<group_mapped> (group_mapped.hxx:146-235): a
// group of TPT lanes owns TPT rows and visits their nonzeros TPT at a time;
// each chunk is reduced by row with a segmented scan and added to per-row
// LDS accumulators (instead of an atomicAdd per nonzero).
template <int TPB, int TPT, typename S>
__global__ __launch_bounds__(TPB) void typed_group_mapped_kernel(
    const int* __restrict__ offsets, const int* __restrict__ indices,
    const S* __restrict__ values, const S* __restrict__ x, S* __restrict__ y, int rows,
    int nnz) {
  using setup_t =
      loops::schedule::setup<algorithms_t::group_mapped, TPB, TPT, int, int, long long, long long>;
  __shared__ typename setup_t::storage_t storage;
  __shared__ S s_acc[TPB];
  __shared__ int s_key[TPB];
  __shared__ S s_wv[TPB / 64 > 0 ? TPB / 64 : 1];
  __shared__ int s_wf[TPB / 64 > 0 ? TPB / 64 : 1];
  setup_t config(storage, const_cast<int*>(offsets), rows, nnz);
  auto p = config.partition();
  auto atoms = config.atom_accessor(p);
  const int aggregate = *atoms.end();
  const int rank = (int)p.thread_rank();
  const int gbase = (int)p.meta_group_rank() * TPT;
  s_acc[threadIdx.x] = S(0);
  p.sync();
  int trip = aggregate;
  if constexpr (TPT > 64 && TPT < TPB) {
    for (int g = 0; g < TPB / TPT; ++g)
      trip = storage.tile_aggregates[g] > trip ? storage.tile_aggregates[g] : trip;
  }
  for (int base = 0; base < trip; base += TPT) {
    const int v_atom = base + rank;
    int key = (TPT > 64 ? TPB : TPT) + (TPT > 64 ? (int)p.meta_group_rank() : 0);
    S prod = S(0);
    if (v_atom < aggregate) {
      int va = v_atom;
      int v_tile = config.tile_accessor(va, p);
      int row = config.tile_id(v_tile, p);
      int nz = config.atom_id(va, row, v_tile, p);
      prod = values[nz] * x[indices[nz]];
      key = TPT > 64 ? gbase + v_tile : v_tile;
    }
    S v = prod;
    if constexpr (TPT <= 64) {
      const int kp = __shfl_up(key, 1, TPT), kn = __shfl_down(key, 1, TPT);
      const int head = (rank == 0 || kp != key) ? 1 : 0;
      const int tail = (rank == TPT - 1 || kn != key) ? 1 : 0;
      int f = head;
      typed_seg_scan<TPT>(v, f);
      if (tail && key < TPT) s_acc[gbase + key] = s_acc[gbase + key] + v;
    } else {
      s_key[threadIdx.x] = key;
      __syncthreads();
      const int head = (threadIdx.x == 0 || s_key[threadIdx.x - 1] != key) ? 1 : 0;
      const int tail = (threadIdx.x == TPB - 1 || s_key[threadIdx.x + 1] != key) ? 1 : 0;
      int f = head;
      S pv;
      int pf;
      typed_block_scan<TPB>(v, f, s_wv, s_wf, pv, pf);
      if (tail && key < TPB) s_acc[key] = s_acc[key] + v;
    }
    p.sync();
  }
  if (rank < config.get_length(p)) y[storage.tiles_indices[gbase + rank]] = s_acc[gbase + rank];
}

// thread_mapped through setup<thread_mapped> and the schedule-free original:
// one row per work-item, summed in CSR order.
template <typename S>
__global__ void typed_thread_mapped_kernel(
    loops::schedule::setup<algorithms_t::thread_mapped, 1, 1, int, int, long long, long long>
        config,
    const int* __restrict__ indices, const S* __restrict__ values, const S* __restrict__ x,
    S* __restrict__ y) {
  for (auto row : config.tiles()) {
    S sum = S(0);
    for (auto nz : config.atoms(row)) sum = sum + values[nz] * x[indices[nz]];
    y[row] = sum;
  }
}

template <typename S>
__global__ void typed_original_kernel(const int* __restrict__ offsets,
                                      const int* __restrict__ indices,
                                      const S* __restrict__ values, const S* __restrict__ x,
                                      S* __restrict__ y, int rows) {
  const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  S sum = S(0);
  for (int nz = offsets[row]; nz < offsets[row + 1]; ++nz) sum = sum + values[nz] * x[indices[nz]];
  y[row] = sum;
}

inline hipError_t rec(hipEvent_t e, hipStream_t s) { return e ? hipEventRecord(e, s) : hipSuccess; }

template <typename S>
hipError_t fixup_t(long long units, const typed_fixup<S>& w, S* y, hipStream_t s) {
  if (units <= 0) return hipSuccess;
  typed_fixup_kernel<S><<<(unsigned)((units + 255) / 256), 256, 0, s>>>(units, w, y);
  return hipGetLastError();
}

template <int TPB, int IPT, typename S>
hipError_t mp_t(const csr_args& A, const S* x, S* y, const loops_coord_t* coords, void* ws,
                hipEvent_t ev0, hipEvent_t ev1, hipStream_t s) {
  hipError_t e = rec(ev0, s);
  if (e != hipSuccess) return e;
  const long long total = (long long)A.rows + A.nnz;
  const long long tiles = (total + TPB * IPT - 1) / (TPB * IPT);
  if (tiles == 0) return rec(ev1, s);
  const typed_fixup<S> w = carve_typed_fixup<S>(ws, tiles);
  typed_merge_path_kernel<TPB, IPT, S><<<(unsigned)tiles, TPB, 0, s>>>(
      A.offsets, A.indices, static_cast<const S*>(A.values), x, y, A.rows, A.nnz, coords, w);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = rec(ev1, s)) != hipSuccess) return e;
  return tiles > 1 ? fixup_t<S>(tiles, w, y, s) : hipSuccess;
}

template <int TPB, typename S>
hipError_t wo_t(const csr_args& A, const S* x, S* y, int grid, void* ws, hipEvent_t ev0,
                hipEvent_t ev1, hipStream_t s) {
  hipError_t e = rec(ev0, s);
  if (e != hipSuccess) return e;
  if ((long long)A.rows + A.nnz == 0) return rec(ev1, s);
  const long long units = (long long)grid * TPB;
  const typed_fixup<S> w = carve_typed_fixup<S>(ws, units);
  typed_work_oriented_kernel<TPB, S><<<(unsigned)grid, TPB, 0, s>>>(
      A.offsets, A.indices, static_cast<const S*>(A.values), x, y, A.rows, A.nnz, w);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = rec(ev1, s)) != hipSuccess) return e;
  return fixup_t<S>(units, w, y, s);
}

template <int TPB, int TPT, typename S>
hipError_t gm_t(const csr_args& A, const S* x, S* y, hipEvent_t ev0, hipEvent_t ev1,
                hipStream_t s) {
  hipError_t e = rec(ev0, s);
  if (e != hipSuccess) return e;
  const long long grid = ((long long)A.rows + TPB - 1) / TPB;
  if (grid > 0) {
    typed_group_mapped_kernel<TPB, TPT, S><<<(unsigned)grid, TPB, 0, s>>>(
        A.offsets, A.indices, static_cast<const S*>(A.values), x, y, A.rows, A.nnz);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return rec(ev1, s);
}

template <typename S>
hipError_t tm_t(const csr_args& A, const S* x, S* y, int tpb, bool original, hipEvent_t ev0,
                hipEvent_t ev1, hipStream_t s) {
  hipError_t e = rec(ev0, s);
  if (e != hipSuccess) return e;
  const long long grid = ((long long)A.rows + tpb - 1) / tpb;
  if (grid > 0) {
    const S* v = static_cast<const S*>(A.values);
    if (original) {
      typed_original_kernel<S><<<(unsigned)grid, tpb, 0, s>>>(A.offsets, A.indices, v, x, y,
                                                               A.rows);
    } else {
      loops::schedule::setup<algorithms_t::thread_mapped, 1, 1, int, int, long long, long long>
          config(const_cast<int*>(A.offsets), A.rows, A.nnz);
      typed_thread_mapped_kernel<S><<<(unsigned)grid, tpb, 0, s>>>(config, A.indices, v, x, y);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return rec(ev1, s);
}

#define LOOPS_TYPED_MP(X) X(64, 3) X(128, 5) X(256, 4) X(256, 8)
#define LOOPS_TYPED_WO(X) X(64) X(128) X(256)
#define LOOPS_TYPED_GM(X) X(128, 128) X(256, 256) X(256, 128) X(256, 64) X(256, 32) X(64, 64)

template <typename S>
hipError_t launch_t(const typed_launch& L, const csr_args& A, const void* xv, void* yv,
                    hipStream_t s) {
  const S* x = static_cast<const S*>(xv);
  S* y = static_cast<S*>(yv);
  switch (L.schedule) {
    case LOOPS_SCHEDULE_MERGE_PATH_FLAT:
#define X(T_, I_) \
  if (L.tpb == T_ && L.ipt == I_) return mp_t<T_, I_, S>(A, x, y, L.coords, L.ws, L.ev0, L.ev1, s);
      LOOPS_TYPED_MP(X)
#undef X
      break;
    case LOOPS_SCHEDULE_WORK_ORIENTED:
#define X(T_) \
  if (L.tpb == T_) return wo_t<T_, S>(A, x, y, L.grid, L.ws, L.ev0, L.ev1, s);
      LOOPS_TYPED_WO(X)
#undef X
      break;
    case LOOPS_SCHEDULE_GROUP_MAPPED:
#define X(T_, G_) \
  if (L.tpb == T_ && L.tpt == G_) return gm_t<T_, G_, S>(A, x, y, L.ev0, L.ev1, s);
      LOOPS_TYPED_GM(X)
#undef X
      break;
    case LOOPS_SCHEDULE_THREAD_MAPPED:
      return tm_t<S>(A, x, y, L.tpb, false, L.ev0, L.ev1, s);
    case LOOPS_SCHEDULE_ORIGINAL:
      return tm_t<S>(A, x, y, L.tpb, true, L.ev0, L.ev1, s);
    default:
      break;
  }
  return hipErrorInvalidValue;
}

}  // namespace

bool typed_spmv_supported(loops_schedule_t schedule, int tpb, int ipt, int tpt) {
  switch (schedule) {
    case LOOPS_SCHEDULE_MERGE_PATH_FLAT:
#define X(T_, I_) \
  if (tpb == T_ && ipt == I_) return true;
      LOOPS_TYPED_MP(X)
#undef X
      return false;
    case LOOPS_SCHEDULE_WORK_ORIENTED:
#define X(T_) \
  if (tpb == T_) return true;
      LOOPS_TYPED_WO(X)
#undef X
      return false;
    case LOOPS_SCHEDULE_GROUP_MAPPED:
#define X(T_, G_) \
  if (tpb == T_ && tpt == G_) return true;
      LOOPS_TYPED_GM(X)
#undef X
      return false;
    case LOOPS_SCHEDULE_THREAD_MAPPED:
    case LOOPS_SCHEDULE_ORIGINAL:
      return tpb > 0 && tpb <= 1024 && tpb % 64 == 0;
    default:
      return false;
  }
}

hipError_t launch_typed_spmv(const typed_launch& L, const csr_args& A, const void* x, void* y,
                             hipStream_t s) {
  if (A.vtype == LOOPS_VALUE_F64) return launch_t<double>(L, A, x, y, s);
  if (A.vtype == LOOPS_VALUE_I32) return launch_t<uint32_t>(L, A, x, y, s);
  return hipErrorInvalidValue;
}

}  // namespace loops_mi355x
