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
//   SPMM   256/G merge-path units of an SpMM item, G lanes spanning its
//          columns per unit (the single-matrix SpMM merge-path unit).
// The caller copies the table to the device and calls loops_batch_launch:
// a persistent kernel (a few workgroups per CU) strides over the work list,
// then the two fix-up kernels finish rows that cross merge tiles / units.
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

enum batch_kind : int { BK_MERGE = 0, BK_ROWS = 1, BK_SPMM = 2 };

struct batch_item_dev {
  const int* offsets;
  const int* indices;
  const void* values;
  const float* x;
  float* y;
  fixup_ws fx;  // MERGE: this item's tile slots
  spmm_ws sx;   // SPMM: this item's unit slots
  long long ldx, ldy;
  int rows, cols, nnz, k;
  int vtype, kind, ipt, g;  // ipt: merge steps per lane (MERGE), per unit (SPMM)
  int units, first_slot, pad0, pad1;
};

struct batch_work {
  int item, kind, sub, pad;
};

struct batch_header {
  long long n_items, n_work, n_tiles, n_units;
  long long items_off, work_off, tile_item_off, unit_item_off;
};

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
constexpr int kBatchTpb = 256;
using batch_smem = mp_smem<kBatchTpb, 16>;

template <int IPT, typename V>
__device__ __forceinline__ void run_merge(const batch_item_dev& it, int tile,
                                          batch_smem& raw) {
  auto& sm = *reinterpret_cast<mp_smem<kBatchTpb, IPT>*>(&raw);
  mp_tile<kBatchTpb, IPT, V, true>(it.offsets, it.indices,
                                   static_cast<const V*>(it.values), it.x, it.y,
                                   it.rows, it.nnz, tile, nullptr, it.fx, sm);
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

template <int G, typename V>
__device__ __forceinline__ void run_spmm(const batch_item_dev& it, int first) {
  const long long u = first + (long long)(threadIdx.x / G);
  if (u >= it.units) return;  // uniform within the group
  spmm_merge_unit<G, 1, V>(it.offsets, it.indices, static_cast<const V*>(it.values),
                           it.rows, it.nnz, it.k, it.x, it.ldx, it.y, it.ldy,
                           it.ipt, u, it.sx);
}

template <typename V>
__device__ __forceinline__ void run_item(const batch_item_dev& it,
                                         const batch_work& w, batch_smem& sm) {
  switch (w.kind) {
    case BK_MERGE:
      if (it.ipt == 4) run_merge<4, V>(it, w.sub, sm);
      else if (it.ipt == 8) run_merge<8, V>(it, w.sub, sm);
      else run_merge<16, V>(it, w.sub, sm);
      break;
    case BK_ROWS:
      run_rows<V>(it, w.sub);
      break;
    case BK_SPMM:
      switch (it.g) {
        case 1: run_spmm<1, V>(it, w.sub); break;
        case 2: run_spmm<2, V>(it, w.sub); break;
        case 4: run_spmm<4, V>(it, w.sub); break;
        case 8: run_spmm<8, V>(it, w.sub); break;
        case 16: run_spmm<16, V>(it, w.sub); break;
        case 32: run_spmm<32, V>(it, w.sub); break;
        default: run_spmm<64, V>(it, w.sub); break;
      }
      break;
  }
}

__global__ __launch_bounds__(kBatchTpb) void batch_kernel(const char* __restrict__ table) {
  __shared__ batch_smem sm;
  const batch_header* h = reinterpret_cast<const batch_header*>(table);
  const batch_item_dev* items =
      reinterpret_cast<const batch_item_dev*>(table + h->items_off);
  const batch_work* work = reinterpret_cast<const batch_work*>(table + h->work_off);
  const long long n = h->n_work;
  for (long long i = blockIdx.x; i < n; i += gridDim.x) {
    const batch_work w = work[i];
    const batch_item_dev& it = items[w.item];
    if (it.vtype == LOOPS_VALUE_BF16) run_item<bf16_t>(it, w, sm);
    else run_item<float>(it, w, sm);
    __syncthreads();  // LDS is reused by the next piece
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

// SpMM unit fix-up across the batch: one workgroup per global unit slot.
__global__ __launch_bounds__(64) void batch_spmm_fixup_kernel(const char* __restrict__ table) {
  const batch_header* h = reinterpret_cast<const batch_header*>(table);
  const batch_item_dev* items =
      reinterpret_cast<const batch_item_dev*>(table + h->items_off);
  const int2* unit_item = reinterpret_cast<const int2*>(table + h->unit_item_off);
  const long long s = blockIdx.x;
  if (s >= h->n_units) return;
  const int2 iu = unit_item[s];
  const batch_item_dev& it = items[iu.x];
  spmm_fixup_unit(it.offsets, it.rows, it.k, it.ipt, iu.y, it.sx, it.y, it.ldy);
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
}  // namespace

// Sizes of the table and of the device workspace for these items, and the
// per-item kind / pieces (written to shape[] when not null).
loops_status_t batch_layout(const loops_batch_item_t* items, int64_t n,
                            loops_batch_info_t* info, batch_item_dev* out_items,
                            char* dev_ws) {
  long long n_work = 0, n_tiles = 0, n_units = 0;
  size_t spmv_ws = 0, spmm_ws_bytes = 0;
  // First pass: counts and workspace sizes.
  for (int64_t i = 0; i < n; ++i) {
    const loops_batch_item_t& b = items[i];
    const loops_csr_view_t& A = b.A;
    if (A.rows < 0 || A.cols < 0 || A.nnzs < 0 || b.k < 1 || b.k > (1 << 20))
      return LOOPS_ERROR_INVALID_VALUE;
    if (A.rows + A.nnzs >= (int64_t)INT32_MAX || A.cols >= (int64_t)INT32_MAX)
      return LOOPS_ERROR_OVERFLOW;
    if (A.value_type != LOOPS_VALUE_F32 && A.value_type != LOOPS_VALUE_BF16)
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
      n_work += (A.rows + kBatchTpb - 1) / kBatchTpb;
    } else if (b.k == 1) {
      if (b.schedule != LOOPS_SCHEDULE_MERGE_PATH_FLAT) return LOOPS_ERROR_NOT_SUPPORTED;
      const int ipt = b.items_per_thread ? b.items_per_thread : 8;
      if (ipt != 4 && ipt != 8 && ipt != 16) return LOOPS_ERROR_NOT_SUPPORTED;
      const long long T = (total + (long long)kBatchTpb * ipt - 1) / ((long long)kBatchTpb * ipt);
      n_work += T;
      n_tiles += T;
    } else {
      if (b.schedule != LOOPS_SCHEDULE_MERGE_PATH_FLAT) return LOOPS_ERROR_NOT_SUPPORTED;
      const long long ipu = b.items_per_thread ? b.items_per_thread : 256;
      if (ipu < 16 || ipu > 65536) return LOOPS_ERROR_NOT_SUPPORTED;
      const long long U = (total + ipu - 1) / ipu;
      const int g = next_pow2_le64(b.k);
      n_work += (U + kBatchTpb / g - 1) / (kBatchTpb / g);
      n_units += U;
      spmm_ws_bytes += U > 0 ? a256(spmm_ws_bytes_impl(U, (int)b.k)) : 0;
    }
  }
  spmv_ws = n_tiles > 0 ? fixup_ws_bytes(n_tiles) : 0;
  batch_header hd{};
  hd.n_items = n;
  hd.n_work = n_work;
  hd.n_tiles = n_tiles;
  hd.n_units = n_units;
  hd.items_off = a256(sizeof(batch_header));
  hd.work_off = hd.items_off + a256((size_t)n * sizeof(batch_item_dev));
  hd.tile_item_off = hd.work_off + a256((size_t)n_work * sizeof(batch_work));
  hd.unit_item_off = hd.tile_item_off + a256((size_t)n_tiles * sizeof(int));
  info->table_bytes = (int64_t)(hd.unit_item_off + a256((size_t)n_units * 2 * sizeof(int)));
  info->workspace_bytes = (int64_t)(spmv_ws + spmm_ws_bytes + 256);
  info->work_items = n_work;
  info->spmv_tiles = n_tiles;
  info->spmm_units = n_units;
  if (!out_items) return LOOPS_SUCCESS;

  // Second pass: fill the table (out_items points at its start).
  char* table = reinterpret_cast<char*>(out_items);
  std::memcpy(table, &hd, sizeof(hd));
  batch_item_dev* its = reinterpret_cast<batch_item_dev*>(table + hd.items_off);
  batch_work* work = reinterpret_cast<batch_work*>(table + hd.work_off);
  int* tile_item = reinterpret_cast<int*>(table + hd.tile_item_off);
  int* unit_item = reinterpret_cast<int*>(table + hd.unit_item_off);
  fixup_ws all_tiles = n_tiles > 0 ? carve_fixup_ws(dev_ws, n_tiles) : fixup_ws{};
  char* spmm_base = dev_ws + spmv_ws;
  long long w = 0, slot = 0, unit = 0;
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
      for (long long q = 0; q < nb; ++q) work[w++] = batch_work{(int)i, BK_ROWS, (int)q, 0};
    } else if (b.k == 1) {
      d.kind = BK_MERGE;
      d.ipt = b.items_per_thread ? b.items_per_thread : 8;
      const long long T = (total + (long long)kBatchTpb * d.ipt - 1) / ((long long)kBatchTpb * d.ipt);
      d.first_slot = (int)slot;
      d.fx.carry_row = all_tiles.carry_row + slot;
      d.fx.seg_end = all_tiles.seg_end + slot;
      d.fx.carry_val = all_tiles.carry_val + slot;
      d.fx.head_val = all_tiles.head_val + slot;
      for (long long t = 0; t < T; ++t) {
        work[w++] = batch_work{(int)i, BK_MERGE, (int)t, 0};
        tile_item[slot + t] = (int)i;
      }
      slot += T;
    } else {
      d.kind = BK_SPMM;
      d.ipt = b.items_per_thread ? b.items_per_thread : 256;
      d.g = next_pow2_le64(b.k);
      const long long U = (total + d.ipt - 1) / d.ipt;
      d.units = (int)U;
      if (U > 0) {
        d.sx = carve_spmm_ws(spmm_base, U, (int)b.k);
        spmm_base += a256(spmm_ws_bytes_impl(U, (int)b.k));
      }
      const int per = kBatchTpb / d.g;
      for (long long q = 0; q < U; q += per) work[w++] = batch_work{(int)i, BK_SPMM, (int)q, 0};
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

hipError_t batch_launch(const void* dev_table, const loops_batch_info_t* info,
                        int grid, hipStream_t s) {
  const char* t = static_cast<const char*>(dev_table);
  if (info->work_items > 0) {
    long long g = grid > 0 ? grid : 1024;
    if (g > info->work_items) g = info->work_items;
    batch_kernel<<<(unsigned)g, kBatchTpb, 0, s>>>(t);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (info->spmv_tiles > 0) {
    batch_spmv_fixup_kernel<<<(unsigned)((info->spmv_tiles + 255) / 256), 256, 0, s>>>(t);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (info->spmm_units > 0) {
    batch_spmm_fixup_kernel<<<(unsigned)info->spmm_units, 64, 0, s>>>(t);
    return hipGetLastError();
  }
  return hipSuccess;
}

}  // namespace loops_mi355x
