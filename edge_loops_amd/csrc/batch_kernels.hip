// Batches of independent products (many matrices, SpMV and multi-RHS SpMM
// mixed) in one persistent kernel.
//
// The host fills a table (loops_batch_prepare): one descriptor per item
// (pointers, sizes, the schedule chosen for it) and a work list of
// workgroup-sized pieces:
//   MERGE  one merge tile of an SpMV item (256 lanes x ipt steps, ipt 4/8/16),
//          the same tile code as the single-matrix merge-path kernel;
//   ROWS   256 consecutive rows of an SpMV item, one row per lane (the
//          thread_mapped schedule) for short, even rows;
//   GROUP  256/G consecutive rows of an SpMV item, G lanes per row (G = 2..64,
//          a power of two) with a fixed xor-butterfly reduction, for rows
//          of moderate, even length;
//   SPMM   one merge-path SpMM tile of an SpMM item (the single-matrix SpMM
//          tile: ranges of G lanes spanning its columns, 1 or 4 columns per
//          lane, IPT = 4, 8 or 16 steps per range).
// The caller copies the table to the device and calls loops_batch_launch:
// a prologue kernel zeroes the work-queue counters and writes every merge
// tile's start coordinate (one lane per coordinate), a persistent SpMV
// kernel works through the SpMV pieces and a persistent SpMM kernel through
// the SpMM pieces (separate kernels so that each keeps its own register
// budget and occupancy), then the two fix-up kernels finish rows that cross
// merge tiles / units.  Persistent workgroups take pieces from a device-wide
// counter, so uneven pieces balance dynamically.
// Every item is computed exactly as it would be alone with the same
// schedule and tile size, so its result does not depend on what else is in
// the batch.  Launching allocates nothing and can be captured in a hipGraph.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "device_common.hpp"
#include "loops_internal.hpp"
#include "merge_tile.hpp"
#include "spmm_unit.hpp"

namespace loops_mi355x {

enum batch_kind : int { BK_MERGE = 0, BK_ROWS = 1, BK_SPMM = 2, BK_GROUP = 3 };

struct batch_item_dev {
  const int* offsets;
  const int* indices;
  const void* values;
  const float* x;
  float* y;
  fixup_ws fx;  // MERGE: this item's tile slots
  loops_coord_t* coords;  // MERGE, SPMM: tiles + 1 start coordinates
  spmm_ws sx;   // SPMM: this item's tile slots
  long long ldx, ldy;
  int rows, cols, nnz, k;
  int vtype, kind, ipt, g;  // ipt: merge steps per lane (MERGE) / range (SPMM)
                            // g: lanes per row (GROUP) or per range (SPMM)
  int units, first_slot, cpl, first_coord;  // units: SPMM tiles; cpl: columns per lane
  int ti, pad0, pad1, pad2;                 // ti: merge steps per tile (MERGE, SPMM)
};

struct batch_work {
  int item, kind, sub, pad;
};

// Work list: SpMV pieces in [0, n_spmv_work), SpMM pieces after them.
struct batch_header {
  long long n_items, n_work, n_spmv_work, n_tiles, n_coords, n_units;
  long long items_off, work_off, tile_item_off, coord_item_off, unit_item_off;
  unsigned long long* queue;  // work-queue counters: SpMV kernel, then SpMM
};

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
constexpr int kBatchTpb = 256;

// Work-queue counters per persistent kernel.  Pieces are split into nq
// interleaved subsets (piece i in subset i % nq), each with its own counter
// on its own 256-byte line, so that claims spread over memory channels
// instead of serialising on one address.
#ifndef LOOPS_BATCH_QUEUES
#define LOOPS_BATCH_QUEUES 8
#endif
constexpr int kQueues = LOOPS_BATCH_QUEUES;
constexpr int kQueueStride = 32;  // unsigned long longs between counters
constexpr size_t kQueueBytes = 2 * kQueues * kQueueStride * sizeof(unsigned long long);

// Next piece of workgroup blockIdx.x: subset q = blockIdx.x % nq holds
// pieces q, q + nq, ...; its wq workgroups start with the first wq of them
// and claim the rest through counter q.
struct piece_queue {
  unsigned long long* ctr;
  long long nq, q, wq;
  __device__ piece_queue(unsigned long long* base) {
    nq = gridDim.x < kQueues ? gridDim.x : kQueues;
    q = blockIdx.x % nq;
    wq = ((long long)gridDim.x - q + nq - 1) / nq;
    ctr = base + q * kQueueStride;
  }
  __device__ long long first() const { return q + nq * (long long)(blockIdx.x / nq); }
  __device__ long long claim() const {
    return q + nq * (wq + (long long)atomicAdd(ctr, 1ULL));
  }
};

// The SpMV kernel is instantiated for the largest merge tile in the batch
// (MAXIPT = 4, 8 or 16 steps per lane): its LDS and registers are sized for
// that tile, so batches of smaller tiles keep more workgroups resident.
template <int IPT, typename V, int MAXIPT>
__device__ __forceinline__ void run_merge(const batch_item_dev& it, int tile,
                                          mp_smem<kBatchTpb, MAXIPT>& raw) {
  static_assert(IPT <= MAXIPT, "tile larger than the kernel's LDS");
  auto& sm = *reinterpret_cast<mp_smem<kBatchTpb, IPT>*>(&raw);
  mp_tile<kBatchTpb, IPT, V, true>(it.offsets, it.indices,
                                   static_cast<const V*>(it.values), it.x, it.y,
                                   it.rows, it.nnz, tile, it.coords, it.fx, sm);
}

template <typename V>
__device__ __forceinline__ void run_rows(const batch_item_dev& it, int block) {
  const long long row = (long long)block * kBatchTpb + threadIdx.x;
  if (row >= it.rows) return;
  const V* vals = static_cast<const V*>(it.values);
  float sum = 0.f;
  for (int nz = it.offsets[row]; nz < it.offsets[row + 1]; ++nz)
    sum += value_traits<V>::load1(vals + nz) * it.x[ld_idx(it.indices + nz)];
  it.y[row] = sum;
}

// G lanes per row; lane l sums nonzeros l, l+G, ... of the row, then the
// group reduces with an xor butterfly (a fixed order).
template <int G, typename V>
__device__ __forceinline__ void run_group(const batch_item_dev& it, int block) {
  const long long row = (long long)block * (kBatchTpb / G) + threadIdx.x / G;
  const int lane = threadIdx.x & (G - 1);
  float sum = 0.f;
  if (row < it.rows) {
    const V* vals = static_cast<const V*>(it.values);
    const int e = it.offsets[row + 1];
    for (int nz = it.offsets[row] + lane; nz < e; nz += G)
      sum += value_traits<V>::load1(vals + nz) * it.x[ld_idx(it.indices + nz)];
  }
#pragma unroll
  for (int m = G / 2; m >= 1; m >>= 1) sum += __shfl_xor(sum, m, G);
  if (row < it.rows && lane == 0) it.y[row] = sum;
}

using batch_spmm_smem = spmm_smem<16, 4>;  // the largest SpMM tile's LDS

template <int IPT, int CPL, typename V>
__device__ __forceinline__ void run_spmm(const batch_item_dev& it, int tile,
                                         batch_spmm_smem& raw) {
  auto& sm = *reinterpret_cast<spmm_smem<IPT, CPL>*>(&raw);
  spmm_tile<IPT, CPL, V>(it.offsets, it.indices, static_cast<const V*>(it.values), it.rows,
                         it.nnz, it.k, it.x, it.ldx, it.y, it.ldy, it.g, tile, it.coords,
                         it.sx, sm);
}

template <typename V, int MAXIPT>
__device__ __forceinline__ void run_spmv_piece(const batch_item_dev& it,
                                               const batch_work& w,
                                               mp_smem<kBatchTpb, MAXIPT>& sm) {
  switch (w.kind) {
    case BK_MERGE:
      if constexpr (MAXIPT >= 16) {
        if (it.ipt == 16) { run_merge<16, V, MAXIPT>(it, w.sub, sm); break; }
      }
      if constexpr (MAXIPT >= 8) {
        if (it.ipt == 8) { run_merge<8, V, MAXIPT>(it, w.sub, sm); break; }
      }
      run_merge<4, V, MAXIPT>(it, w.sub, sm);
      break;
    case BK_ROWS:
      run_rows<V>(it, w.sub);
      break;
    case BK_GROUP:
      switch (it.g) {
        case 2: run_group<2, V>(it, w.sub); break;
        case 4: run_group<4, V>(it, w.sub); break;
        case 8: run_group<8, V>(it, w.sub); break;
        case 16: run_group<16, V>(it, w.sub); break;
        case 32: run_group<32, V>(it, w.sub); break;
        default: run_group<64, V>(it, w.sub); break;
      }
      break;
  }
}

template <typename V>
__device__ __forceinline__ void run_spmm_piece(const batch_item_dev& it, const batch_work& w,
                                               batch_spmm_smem& sm) {
  if (it.cpl == 4) {
    if (it.ipt == 4) run_spmm<4, 4, V>(it, w.sub, sm);
    else if (it.ipt == 16) run_spmm<16, 4, V>(it, w.sub, sm);
    else run_spmm<8, 4, V>(it, w.sub, sm);
  } else {
    if (it.ipt == 4) run_spmm<4, 1, V>(it, w.sub, sm);
    else if (it.ipt == 16) run_spmm<16, 1, V>(it, w.sub, sm);
    else run_spmm<8, 1, V>(it, w.sub, sm);
  }
}

// Prologue: zero the work-queue counters; write the start coordinate of
// every merge tile of every MERGE and SPMM item (and the end of its last
// tile): coordinate c belongs to item coord_item[c].
__global__ __launch_bounds__(256) void batch_prologue_kernel(const char* __restrict__ table) {
  const batch_header* h = reinterpret_cast<const batch_header*>(table);
  const batch_item_dev* items =
      reinterpret_cast<const batch_item_dev*>(table + h->items_off);
  const int* coord_item = reinterpret_cast<const int*>(table + h->coord_item_off);
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < 2 * kQueues) h->queue[c * kQueueStride] = 0;
  if (c >= h->n_coords) return;
  const batch_item_dev& it = items[coord_item[c]];
  const long long t = c - it.first_coord;
  const long long total = (long long)it.rows + it.nnz;
  long long d = t * (long long)it.ti;
  if (d > total) d = total;
  const auto xy = loops::search::_binary_search<long long>(
      d, it.offsets + 1, 0LL, (long long)it.rows, (long long)it.nnz);
  it.coords[t] = loops_coord_t{xy.x, xy.y};
}

// Persistent kernels (see piece_queue).  Lane 0 claims the next piece as
// the current one starts, so the atomic's latency overlaps the piece; every
// workgroup leaves once its claim reaches n.
template <int MAXIPT>
__global__ __launch_bounds__(kBatchTpb) void batch_spmv_kernel(const char* __restrict__ table) {
  __shared__ mp_smem<kBatchTpb, MAXIPT> sm;
  __shared__ long long s_next;
  const batch_header* h = reinterpret_cast<const batch_header*>(table);
  const batch_item_dev* items =
      reinterpret_cast<const batch_item_dev*>(table + h->items_off);
  const batch_work* work = reinterpret_cast<const batch_work*>(table + h->work_off);
  const long long n = h->n_spmv_work;
  const piece_queue pq(h->queue);
  long long i = pq.first();
  while (i < n) {
    long long claim = 0;
    if (threadIdx.x == 0) claim = pq.claim();
    const batch_work w = work[i];
    const batch_item_dev& it = items[w.item];
    if (it.vtype == LOOPS_VALUE_BF16) run_spmv_piece<bf16_t, MAXIPT>(it, w, sm);
    else if (it.vtype == LOOPS_VALUE_F16) run_spmv_piece<f16_t, MAXIPT>(it, w, sm);
    else run_spmv_piece<float, MAXIPT>(it, w, sm);
    if (threadIdx.x == 0) s_next = claim;
    __syncthreads();
    i = s_next;
    __syncthreads();  // s_next and the piece's LDS are reused
  }
}

__global__ __launch_bounds__(kBatchTpb) void batch_spmm_kernel(const char* __restrict__ table) {
  __shared__ batch_spmm_smem sm;
  __shared__ long long s_next;
  const batch_header* h = reinterpret_cast<const batch_header*>(table);
  const batch_item_dev* items =
      reinterpret_cast<const batch_item_dev*>(table + h->items_off);
  const batch_work* work = reinterpret_cast<const batch_work*>(table + h->work_off);
  const long long n = h->n_work - h->n_spmv_work;
  const piece_queue pq(h->queue + kQueues * kQueueStride);
  long long i = pq.first();
  while (i < n) {
    long long claim = 0;
    if (threadIdx.x == 0) claim = pq.claim();
    const batch_work w = work[h->n_spmv_work + i];
    const batch_item_dev& it = items[w.item];
    if (it.vtype == LOOPS_VALUE_BF16) run_spmm_piece<bf16_t>(it, w, sm);
    else if (it.vtype == LOOPS_VALUE_F16) run_spmm_piece<f16_t>(it, w, sm);
    else run_spmm_piece<float>(it, w, sm);
    if (threadIdx.x == 0) s_next = claim;
    __syncthreads();
    i = s_next;
    __syncthreads();
  }
}

// Merge-tile fix-up across the batch: slot s belongs to item tile_item[s].
__global__ __launch_bounds__(256) void batch_spmv_fixup_kernel(const char* __restrict__ table) {
  const batch_header* h = reinterpret_cast<const batch_header*>(table);
  const batch_item_dev* items =
      reinterpret_cast<const batch_item_dev*>(table + h->items_off);
  const int* tile_item = reinterpret_cast<const int*>(table + h->tile_item_off);
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int item = 0;
  long long t = 0, t1 = -1;
  int r = -1;
  if (s < h->n_tiles) {
    item = tile_item[s];
    t = s - items[item].first_slot;
    t1 = items[item].fx.seg_end[t];
    if (t1 >= 0) r = items[item].fx.carry_row[t];
  }
  if (t1 >= 0 && segment_is_short(t, t1))
    items[item].y[r] =
        short_segment_sum(items[item].fx.carry_val, items[item].fx.head_val, t, t1);
  unsigned long long todo = __ballot(t1 >= 0 && !segment_is_short(t, t1));
  while (todo) {
    const int src = __ffsll((long long)todo) - 1;
    todo &= todo - 1;
    const int it = __shfl(item, src, 64);
    const long long a = __shfl(t, src, 64);
    const long long b = __shfl(t1, src, 64);
    const int rr = __shfl(r, src, 64);
    const float v = wave_segment_sum(items[it].fx.carry_val, items[it].fx.head_val, a, b);
    if (lane == src) items[it].y[rr] = v;
  }
}

// SpMM tile fix-up across the batch: one workgroup per global tile slot.
__global__ __launch_bounds__(64) void batch_spmm_fixup_kernel(const char* __restrict__ table) {
  const batch_header* h = reinterpret_cast<const batch_header*>(table);
  const batch_item_dev* items =
      reinterpret_cast<const batch_item_dev*>(table + h->items_off);
  const int2* unit_item = reinterpret_cast<const int2*>(table + h->unit_item_off);
  const long long s = blockIdx.x;
  if (s >= h->n_units) return;
  const int2 iu = unit_item[s];
  const batch_item_dev& it = items[iu.x];
  spmm_fixup_unit(it.offsets, it.rows, it.k, it.ti, iu.y, it.sx, it.y, it.ldy);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {
inline size_t a256(size_t b) { return (b + 255) & ~size_t(255); }
inline int next_pow2_le64(long long k) {
  int g = 1;
  while (g < 64 && g < k) g *= 2;
  return g;
}
// Lanes per row of a GROUP item: the power of two nearest above the mean row
// length, 2..64, unless given.
inline int group_lanes(const loops_batch_item_t& b) {
  if (b.items_per_thread) return b.items_per_thread;
  const long long mean = b.A.rows > 0 ? (b.A.nnzs + b.A.rows - 1) / b.A.rows : 1;
  const int g = next_pow2_le64(mean);
  return g < 2 ? 2 : g;
}
inline bool valid_group(int g) {
  return g == 2 || g == 4 || g == 8 || g == 16 || g == 32 || g == 64;
}
// SpMM lanes per unit and columns per lane: 4 columns per lane (16-byte
// loads and stores) when k >= 16 and B / C rows are 16-byte aligned.
struct spmm_pick {
  int g, cpl;
};
inline spmm_pick spmm_shape_for(const loops_batch_item_t& b) {
  const bool vec4 = b.k >= 16 && b.ldx % 4 == 0 && b.ldy % 4 == 0 &&
                    reinterpret_cast<uintptr_t>(b.x) % 16 == 0 &&
                    reinterpret_cast<uintptr_t>(b.y) % 16 == 0;
  if (vec4) {
    int g = 4;
    while (g < 64 && g * 4 < b.k) g *= 2;
    return {g, 4};
  }
  const int g = next_pow2_le64(b.k);
  return {g < kSpmmMinG ? kSpmmMinG : g, 1};
}
}  // namespace

// Sizes of the table and of the device workspace for these items; with
// out_table, also fills the table (host memory) pointing into dev_ws.
loops_status_t batch_layout(const loops_batch_item_t* items, int64_t n,
                            loops_batch_info_t* info, batch_item_dev* out_table,
                            char* dev_ws) {
  long long n_spmv_work = 0, n_spmm_work = 0, n_tiles = 0, n_coords = 0, n_units = 0;
  int max_ipt = 0;
  size_t spmm_ws_bytes = 0;
  // First pass: validation, counts and workspace sizes.
  for (int64_t i = 0; i < n; ++i) {
    const loops_batch_item_t& b = items[i];
    const loops_csr_view_t& A = b.A;
    if (A.rows < 0 || A.cols < 0 || A.nnzs < 0 || b.k < 1 || b.k > (1 << 20))
      return LOOPS_ERROR_INVALID_VALUE;
    if (A.rows + A.nnzs >= (int64_t)INT32_MAX || A.cols >= (int64_t)INT32_MAX)
      return LOOPS_ERROR_OVERFLOW;
    if (A.value_type != LOOPS_VALUE_F32 && A.value_type != LOOPS_VALUE_BF16 &&
        A.value_type != LOOPS_VALUE_F16)
      return LOOPS_ERROR_NOT_SUPPORTED;
    if ((A.rows > 0 && (!A.offsets || !b.y)) ||
        (A.nnzs > 0 && (!A.indices || !A.values || !b.x)))
      return LOOPS_ERROR_INVALID_VALUE;
    const uintptr_t va = A.value_type == LOOPS_VALUE_F32 ? 16 : 8;
    if (reinterpret_cast<uintptr_t>(A.indices) % 16 ||
        reinterpret_cast<uintptr_t>(A.values) % va)
      return LOOPS_ERROR_NOT_SUPPORTED;  // batch pieces use 16-byte loads
    if (b.k > 1 && (b.ldx < b.k || b.ldy < b.k)) return LOOPS_ERROR_INVALID_VALUE;
    const long long total = A.rows + A.nnzs;
    if (b.k == 1 && b.schedule == LOOPS_SCHEDULE_THREAD_MAPPED) {
      n_spmv_work += (A.rows + kBatchTpb - 1) / kBatchTpb;
    } else if (b.k == 1 && b.schedule == LOOPS_SCHEDULE_GROUP_MAPPED) {
      const int g = group_lanes(b);
      if (!valid_group(g)) return LOOPS_ERROR_NOT_SUPPORTED;
      n_spmv_work += (A.rows + kBatchTpb / g - 1) / (kBatchTpb / g);
    } else if (b.k == 1) {
      if (b.schedule != LOOPS_SCHEDULE_MERGE_PATH_FLAT) return LOOPS_ERROR_NOT_SUPPORTED;
      const int ipt = b.items_per_thread ? b.items_per_thread : 8;
      if (ipt != 4 && ipt != 8 && ipt != 16) return LOOPS_ERROR_NOT_SUPPORTED;
      const long long T = (total + (long long)kBatchTpb * ipt - 1) / ((long long)kBatchTpb * ipt);
      n_spmv_work += T;
      n_tiles += T;
      n_coords += T + 1;
      if (ipt > max_ipt) max_ipt = ipt;
    } else {
      if (b.schedule != LOOPS_SCHEDULE_MERGE_PATH_FLAT) return LOOPS_ERROR_NOT_SUPPORTED;
      const int ipt = b.items_per_thread ? b.items_per_thread : 8;
      if (ipt != 4 && ipt != 8 && ipt != 16) return LOOPS_ERROR_NOT_SUPPORTED;
      const long long ti = spmm_tile_items(spmm_shape_for(b).g, ipt);
      const long long U = (total + ti - 1) / ti;
      n_spmm_work += U;
      n_units += U;
      n_coords += U + 1;
      spmm_ws_bytes += U > 0 ? a256(spmm_ws_bytes_impl(U, (int)b.k)) : 0;
    }
  }
  if (n_spmv_work + n_spmm_work >= (long long)INT32_MAX || n_tiles >= INT32_MAX ||
      n_coords >= INT32_MAX)
    return LOOPS_ERROR_OVERFLOW;
  const size_t tile_ws = n_tiles > 0 ? a256(fixup_ws_bytes(n_tiles)) : 0;
  const size_t coord_ws = a256((size_t)n_coords * sizeof(loops_coord_t));
  batch_header hd{};
  hd.n_items = n;
  hd.n_work = n_spmv_work + n_spmm_work;
  hd.n_spmv_work = n_spmv_work;
  hd.n_tiles = n_tiles;
  hd.n_coords = n_coords;
  hd.n_units = n_units;
  hd.items_off = a256(sizeof(batch_header));
  hd.work_off = hd.items_off + a256((size_t)n * sizeof(batch_item_dev));
  hd.tile_item_off = hd.work_off + a256((size_t)hd.n_work * sizeof(batch_work));
  hd.coord_item_off = hd.tile_item_off + a256((size_t)n_tiles * sizeof(int));
  hd.unit_item_off = hd.coord_item_off + a256((size_t)n_coords * sizeof(int));
  info->table_bytes = (int64_t)(hd.unit_item_off + a256((size_t)n_units * 2 * sizeof(int)));
  info->workspace_bytes = (int64_t)(kQueueBytes + tile_ws + coord_ws + spmm_ws_bytes + 256);
  info->work_items = hd.n_work;
  info->spmv_tiles = n_tiles;
  info->spmm_units = n_units;
  info->spmm_work_items = n_spmm_work;
  info->coordinates = n_coords;
  info->max_items_per_thread = max_ipt;
  if (!out_table) return LOOPS_SUCCESS;

  // Second pass: fill the table.
  char* table = reinterpret_cast<char*>(out_table);
  hd.queue = reinterpret_cast<unsigned long long*>(dev_ws);
  dev_ws += kQueueBytes;
  std::memcpy(table, &hd, sizeof(hd));
  batch_item_dev* its = reinterpret_cast<batch_item_dev*>(table + hd.items_off);
  batch_work* work = reinterpret_cast<batch_work*>(table + hd.work_off);
  int* tile_item = reinterpret_cast<int*>(table + hd.tile_item_off);
  int* coord_item = reinterpret_cast<int*>(table + hd.coord_item_off);
  int* unit_item = reinterpret_cast<int*>(table + hd.unit_item_off);
  fixup_ws all_tiles = n_tiles > 0 ? carve_fixup_ws(dev_ws, n_tiles) : fixup_ws{};
  loops_coord_t* all_coords = reinterpret_cast<loops_coord_t*>(dev_ws + tile_ws);
  char* spmm_base = dev_ws + tile_ws + coord_ws;
  long long wv = 0, wm = n_spmv_work, slot = 0, coord = 0, unit = 0;
  for (int64_t i = 0; i < n; ++i) {
    const loops_batch_item_t& b = items[i];
    batch_item_dev d{};
    d.offsets = b.A.offsets;
    d.indices = b.A.indices;
    d.values = b.A.values;
    d.x = b.x;
    d.y = b.y;
    d.ldx = b.k > 1 ? b.ldx : 1;
    d.ldy = b.k > 1 ? b.ldy : 1;
    d.rows = (int)b.A.rows;
    d.cols = (int)b.A.cols;
    d.nnz = (int)b.A.nnzs;
    d.k = (int)b.k;
    d.vtype = b.A.value_type;
    const long long total = b.A.rows + b.A.nnzs;
    if (b.k == 1 && b.schedule == LOOPS_SCHEDULE_THREAD_MAPPED) {
      d.kind = BK_ROWS;
      const long long nb = (b.A.rows + kBatchTpb - 1) / kBatchTpb;
      for (long long q = 0; q < nb; ++q) work[wv++] = batch_work{(int)i, BK_ROWS, (int)q, 0};
    } else if (b.k == 1 && b.schedule == LOOPS_SCHEDULE_GROUP_MAPPED) {
      d.kind = BK_GROUP;
      d.g = group_lanes(b);
      const long long per = kBatchTpb / d.g;
      const long long nb = (b.A.rows + per - 1) / per;
      for (long long q = 0; q < nb; ++q) work[wv++] = batch_work{(int)i, BK_GROUP, (int)q, 0};
    } else if (b.k == 1) {
      d.kind = BK_MERGE;
      d.ipt = b.items_per_thread ? b.items_per_thread : 8;
      d.ti = kBatchTpb * d.ipt;
      const long long T = (total + d.ti - 1) / d.ti;
      d.first_slot = (int)slot;
      d.fx.carry_row = all_tiles.carry_row + slot;
      d.fx.seg_end = all_tiles.seg_end + slot;
      d.fx.carry_val = all_tiles.carry_val + slot;
      d.fx.head_val = all_tiles.head_val + slot;
      d.first_coord = (int)coord;
      d.coords = all_coords + coord;
      for (long long t = 0; t < T; ++t) {
        work[wv++] = batch_work{(int)i, BK_MERGE, (int)t, 0};
        tile_item[slot + t] = (int)i;
      }
      for (long long t = 0; t <= T; ++t) coord_item[coord + t] = (int)i;
      slot += T;
      coord += T + 1;
    } else {
      d.kind = BK_SPMM;
      d.ipt = b.items_per_thread ? b.items_per_thread : 8;
      const spmm_pick sp = spmm_shape_for(b);
      d.g = sp.g;
      d.cpl = sp.cpl;
      d.ti = (int)spmm_tile_items(d.g, d.ipt);
      const long long U = (total + d.ti - 1) / d.ti;
      d.units = (int)U;
      if (U > 0) {
        d.sx = carve_spmm_ws(spmm_base, U, (int)b.k);
        spmm_base += a256(spmm_ws_bytes_impl(U, (int)b.k));
      }
      d.first_coord = (int)coord;
      d.coords = all_coords + coord;
      for (long long t = 0; t <= U; ++t) coord_item[coord + t] = (int)i;
      coord += U + 1;
      for (long long q = 0; q < U; ++q) work[wm++] = batch_work{(int)i, BK_SPMM, (int)q, 0};
      for (long long q = 0; q < U; ++q) {
        unit_item[2 * (unit + q)] = (int)i;
        unit_item[2 * (unit + q) + 1] = (int)q;
      }
      unit += U;
    }
    its[i] = d;
  }
  return LOOPS_SUCCESS;
}

namespace {
// Resident workgroups of each persistent kernel on a device (occupancy x CUs),
// computed once per device.
struct batch_grids {
  int spmv4 = 0, spmv8 = 0, spmv16 = 0, spmm = 0;
};
constexpr int kMaxDevices = 64;
batch_grids g_grids[kMaxDevices];

hipError_t resident_grids(batch_grids* out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev >= 0 && dev < kMaxDevices && g_grids[dev].spmm > 0) {
    *out = g_grids[dev];
    return hipSuccess;
  }
  int cus = 0, a4 = 0, a8 = 0, a16 = 0, b = 0;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a4, batch_spmv_kernel<4>, kBatchTpb, 0);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a8, batch_spmv_kernel<8>, kBatchTpb, 0);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a16, batch_spmv_kernel<16>, kBatchTpb, 0);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, batch_spmm_kernel, kBatchTpb, 0);
  if (e != hipSuccess) return e;
  batch_grids g;
  g.spmv4 = (a4 > 0 ? a4 : 1) * cus;
  g.spmv8 = (a8 > 0 ? a8 : 1) * cus;
  g.spmv16 = (a16 > 0 ? a16 : 1) * cus;
  g.spmm = (b > 0 ? b : 1) * cus;
  if (dev >= 0 && dev < kMaxDevices) g_grids[dev] = g;
  *out = g;
  return hipSuccess;
}
}  // namespace

hipError_t batch_warm_up() {
  batch_grids g;
  return resident_grids(&g);
}

hipError_t batch_launch(const void* dev_table, const loops_batch_info_t* info,
                        int grid, hipStream_t s) {
  const char* t = static_cast<const char*>(dev_table);
  batch_grids rg;
  hipError_t e = resident_grids(&rg);
  if (e != hipSuccess) return e;
  const long long spmv_work = info->work_items - info->spmm_work_items;
  if (info->work_items > 0) {
    const long long pg = info->coordinates > 0 ? (info->coordinates + 255) / 256 : 1;
    static_assert(2 * kQueues <= 256, "prologue's first block zeroes the counters");
    batch_prologue_kernel<<<(unsigned)pg, 256, 0, s>>>(t);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (spmv_work > 0) {
    const long long mi = info->max_items_per_thread;
    long long g = grid > 0 ? grid : (mi > 8 ? rg.spmv16 : mi > 4 ? rg.spmv8 : rg.spmv4);
    if (g > spmv_work) g = spmv_work;
    if (mi > 8) batch_spmv_kernel<16><<<(unsigned)g, kBatchTpb, 0, s>>>(t);
    else if (mi > 4) batch_spmv_kernel<8><<<(unsigned)g, kBatchTpb, 0, s>>>(t);
    else batch_spmv_kernel<4><<<(unsigned)g, kBatchTpb, 0, s>>>(t);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (info->spmm_work_items > 0) {
    long long g = grid > 0 ? grid : rg.spmm;
    if (g > info->spmm_work_items) g = info->spmm_work_items;
    batch_spmm_kernel<<<(unsigned)g, kBatchTpb, 0, s>>>(t);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (info->spmv_tiles > 0) {
    batch_spmv_fixup_kernel<<<(unsigned)((info->spmv_tiles + 255) / 256), 256, 0, s>>>(t);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (info->spmm_units > 0) {
    batch_spmm_fixup_kernel<<<(unsigned)info->spmm_units, 64, 0, s>>>(t);
    return hipGetLastError();
  }
  return hipSuccess;
}

}  // namespace loops_mi355x
