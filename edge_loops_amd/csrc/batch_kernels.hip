// Batches of independent products (many matrices, SpMV and multi-RHS SpMM
// mixed) in one persistent kernel.
//
// The host fills a table (loops_batch_prepare) whose size grows with the
// number of items, not with their sizes: one descriptor per item (pointers,
// sizes, the schedule chosen for it) and, per kind of piece, the prefix
// counts of each item's pieces.  The pieces are workgroup-sized:
//   MERGE  one merge tile of an SpMV item (256 lanes x ipt steps, ipt 4/8/16),
//          the same tile code as the single-matrix merge-path kernel;
//   ROWS   256 consecutive rows of an SpMV item, one row per lane (the
//          thread_mapped schedule) for short, even rows;
//   GROUP  256/G consecutive rows of an SpMV item, G lanes per row (G = 2..64,
//          a power of two) with a fixed xor-butterfly reduction, for rows
//          of moderate, even length;
//   SPMM   one merge-path SpMM tile of an SpMM item (the single-matrix SpMM
//          tile: ranges of G lanes spanning its columns, 1 or 4 columns per
//          lane, IPT = 4, 8 or 16 steps per range);
//   BLOCKS one (block, segment) stream of an item carrying a row-block
//          plan (the plan's own stream code, 512 lanes, y of the block in
//          LDS); a combine kernel then adds each row's segment partials.
// The pieces of all items form one concatenated range per kernel; a piece
// finds its item by a search over the prefix counts (a 64-lane cooperative
// search, done again only when a workgroup's next piece leaves the item of
// its previous one).
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

#include <algorithm>
#include <mutex>

#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "blk_stream.hpp"
#include "device_common.hpp"
#include "loops_internal.hpp"
#include "merge_tile.hpp"
#include "spmm_unit.hpp"

namespace loops_mi355x {

enum batch_kind : int { BK_MERGE = 0, BK_ROWS = 1, BK_SPMM = 2, BK_GROUP = 3, BK_BLOCKS = 4 };

// Kernel classes, one persistent kernel each, so that every kernel has the
// registers, LDS and occupancy of the single-matrix kernel for its tile:
//   SpMV classes 0..2: merge tiles of 4 (and the ROWS / GROUP pieces), 8 and
//                      16 steps per lane;
//   SpMM variants:     value type x columns per lane (1, 4) x merge steps
//                      per range (4, 8, 16, 32), class kSpmvClasses + variant;
//   row-block streams: class kBlkClass.
constexpr int kSpmvClasses = 3;
constexpr int kSpmmVariants = 3 * 2 * 4;
constexpr int kBlkClass = kSpmvClasses + kSpmmVariants;
constexpr int kClasses = kBlkClass + 1;
// A plan item of several segments has its streams padded to a multiple of
// this many pieces, and such items come first in their class, so that
// their pieces start at multiples of 8 within the class: with 8 queue
// subsets (piece i in subset i % 8, taken by workgroups blockIdx.x % 8 ==
// i % 8, and workgroups go to the 8 XCDs round-robin) segment s of every
// plan then runs on the XCDs s % 8, as in the plan's own launch, and each
// XCD's L2 holds only its slices of x.  One-segment items are not padded.
constexpr int kBlkPiecePad = 8;
__host__ __device__ constexpr int spmm_variant(int vtype, int cpl, int ipt) {
  return (vtype * 2 + (cpl == 4 ? 1 : 0)) * 4 + (ipt == 4 ? 0 : ipt == 8 ? 1 : ipt == 16 ? 2 : 3);
}
__host__ __device__ constexpr int spmv_class(int maxipt) {
  return maxipt == 4 ? 0 : maxipt == 8 ? 1 : 2;
}

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
  int ti, blk, pad1, pad2;  // ti: merge steps per tile (MERGE, SPMM); blk: BLOCKS
                            // item's entry in the table's plan arguments
};

// Prefix arrays of the table, n_items + 1 entries each: entry i = pieces /
// slots / coordinates of items 0 .. i-1.
enum batch_prefix : int {
  BP_TILE = 0,   // merge-tile fix-up slots (MERGE items)
  BP_COORD = 1,  // merge-tile start coordinates (MERGE and SPMM items)
  BP_UNIT = 2,   // SpMM tile fix-up slots (SPMM items)
  BP_BROWS = 3,  // 256-row combine pieces (BLOCKS items)
  BP_COUNT = 4
};

// Items are also listed grouped by kernel class (order, stable in item
// order): class c owns entries [class_begin[c], class_begin[c + 1]) of
// order, and cprefix[j] counts the pieces of entries before j.
struct batch_header {
  long long n_items, n_spmv_work, n_spmm_work, n_tiles, n_coords, n_units;
  long long n_blk_work, n_brows, n_blk;
  long long items_off, prefix_off, order_off, cprefix_off, blk_off;
  long long class_begin[kClasses + 1];
  unsigned long long* queue;  // work-queue counter sets, one per class
};

__device__ __forceinline__ const long long* batch_prefix_of(const char* table,
                                                            const batch_header* h, int kind) {
  return reinterpret_cast<const long long*>(table + h->prefix_off) + kind * (h->n_items + 1);
}

// Item holding index i of a concatenated range: the largest j < n with
// P[j] <= i (P non-decreasing, P[0] = 0 <= i < P[n]), so empty items are
// skipped.  One lane's binary search.
__device__ __forceinline__ int batch_find_item(const long long* __restrict__ P, long long n,
                                               long long i) {
  long long lo = 0, hi = n;  // P[lo] <= i; the answer is in [lo, hi)
  while (hi - lo > 1) {
    const long long mid = (lo + hi) >> 1;
    if (P[mid] <= i) lo = mid;
    else hi = mid;
  }
  return (int)lo;
}

// The same, by a whole wavefront: each round 64 lanes probe 64 evenly spaced
// entries of [lo, hi) and the count of probes <= i narrows the range to
// one step (two rounds for up to 4096 items).  i must be wave-uniform.
__device__ __forceinline__ int batch_wave_find_item(const long long* __restrict__ P,
                                                    long long n, long long i) {
  const int lane = threadIdx.x & 63;
  long long lo = 0, hi = n;
  while (hi - lo > 1) {
    const long long step = (hi - lo + 63) / 64;
    const long long pos = lo + lane * step;
    const bool le = pos < hi && P[pos] <= i;
    const int c = __popcll(__ballot(le));  // >= 1: lane 0 probes lo
    lo += (long long)(c - 1) * step;
    const long long nh = lo + step;
    hi = nh < hi ? nh : hi;
  }
  return (int)lo;
}

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
constexpr int kQueueSets = kClasses;
constexpr size_t kQueueBytes =
    (size_t)kQueueSets * kQueues * kQueueStride * sizeof(unsigned long long);

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

template <typename V, int MAXIPT>
__device__ __forceinline__ void run_spmv_piece(const batch_item_dev& it, int sub,
                                               mp_smem<kBatchTpb, MAXIPT>& sm) {
  switch (it.kind) {
    case BK_MERGE:
      if constexpr (MAXIPT >= 16) {
        if (it.ipt == 16) { run_merge<16, V, MAXIPT>(it, sub, sm); break; }
      }
      if constexpr (MAXIPT >= 8) {
        if (it.ipt == 8) { run_merge<8, V, MAXIPT>(it, sub, sm); break; }
      }
      run_merge<4, V, MAXIPT>(it, sub, sm);
      break;
    case BK_ROWS:
      run_rows<V>(it, sub);
      break;
    case BK_GROUP:
      switch (it.g) {
        case 2: run_group<2, V>(it, sub); break;
        case 4: run_group<4, V>(it, sub); break;
        case 8: run_group<8, V>(it, sub); break;
        case 16: run_group<16, V>(it, sub); break;
        case 32: run_group<32, V>(it, sub); break;
        default: run_group<64, V>(it, sub); break;
      }
      break;
  }
}

// Prologue: zero the work-queue counters; write the start coordinate of
// every merge tile of every MERGE and SPMM item (and the end of its last
// tile).
__global__ __launch_bounds__(256) void batch_prologue_kernel(const char* __restrict__ table) {
  const batch_header* h = reinterpret_cast<const batch_header*>(table);
  const batch_item_dev* items =
      reinterpret_cast<const batch_item_dev*>(table + h->items_off);
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < kQueueSets * kQueues) h->queue[c * kQueueStride] = 0;
  if (c >= h->n_coords) return;
  const long long* P = batch_prefix_of(table, h, BP_COORD);
  const int item = batch_find_item(P, h->n_items, c);
  const batch_item_dev& it = items[item];
  const long long t = c - P[item];
  const long long total = (long long)it.rows + it.nnz;
  long long d = t * (long long)it.ti;
  if (d > total) d = total;
  const auto xy = loops::search::_binary_search<long long>(
      d, it.offsets + 1, 0LL, (long long)it.rows, (long long)it.nnz);
  it.coords[t] = loops_coord_t{xy.x, xy.y};
}

// Persistent kernels (see piece_queue).  Lane 0 claims the next piece as
// the current one starts, so the atomic's latency overlaps the piece; every
// workgroup leaves once its claim reaches n.  Each wavefront keeps the
// piece range [lo, hi) of its current item and searches the prefix counts
// only when a piece falls outside it (the piece index is workgroup-uniform,
// so every wavefront finds the same item).
// The pieces of class cls: workgroups claim them through the class's
// counters and find the item of a piece in the class's prefix counts; fn
// runs one piece (item, index of the piece within the item).
template <typename F>
__device__ __forceinline__ void batch_class_loop(const char* __restrict__ table, int cls,
                                                 long long* s_next, F&& fn) {
  const batch_header* h = reinterpret_cast<const batch_header*>(table);
  const int* order = reinterpret_cast<const int*>(table + h->order_off);
  const long long jb = h->class_begin[cls], je = h->class_begin[cls + 1];
  const long long* P = reinterpret_cast<const long long*>(table + h->cprefix_off) + jb;
  const long long base = P[0], n = P[je - jb] - base;
  const piece_queue pq(h->queue + (long long)cls * kQueues * kQueueStride);
  long long i = pq.first(), lo = 0, hi = 0;
  int item = 0;
  while (i < n) {
    long long claim = 0;
    if (threadIdx.x == 0) claim = pq.claim();
    const long long g = base + i;
    if (g < lo || g >= hi) {
      const int j = batch_wave_find_item(P, je - jb, g);
      item = order[jb + j];
      lo = P[j];
      hi = P[j + 1];
    }
    fn(item, (int)(g - lo));
    if (threadIdx.x == 0) *s_next = claim;
    __syncthreads();
    i = *s_next;
    __syncthreads();  // s_next and the piece's LDS are reused
  }
}

// SpMV class spmv_class(MAXIPT): merge tiles of MAXIPT steps per lane (and,
// in class 0, the ROWS / GROUP pieces).
template <int MAXIPT>
__global__ __launch_bounds__(kBatchTpb) void batch_spmv_kernel(const char* __restrict__ table) {
  __shared__ mp_smem<kBatchTpb, MAXIPT> sm;
  __shared__ long long s_next;
  const batch_header* h = reinterpret_cast<const batch_header*>(table);
  const batch_item_dev* items =
      reinterpret_cast<const batch_item_dev*>(table + h->items_off);
  batch_class_loop(table, spmv_class(MAXIPT), &s_next, [&](int item, int sub) {
    const batch_item_dev& it = items[item];
    if (it.vtype == LOOPS_VALUE_BF16) run_spmv_piece<bf16_t, MAXIPT>(it, sub, sm);
    else if (it.vtype == LOOPS_VALUE_F16) run_spmv_piece<f16_t, MAXIPT>(it, sub, sm);
    else run_spmv_piece<float, MAXIPT>(it, sub, sm);
  });
}

// One instantiation per SpMM tile variant; it works through the pieces of
// that variant's items only, so its registers and LDS are those of the
// single-matrix kernel of the same tile.
template <int IPT, int CPL, typename V>
__global__ __launch_bounds__(kBatchTpb) void batch_spmm_kernel(const char* __restrict__ table) {
  __shared__ spmm_smem<IPT, CPL> sm;
  __shared__ long long s_next;
  constexpr int var = spmm_variant(
      std::is_same<V, float>::value ? LOOPS_VALUE_F32
                                    : std::is_same<V, bf16_t>::value ? LOOPS_VALUE_BF16
                                                                     : LOOPS_VALUE_F16,
      CPL, IPT);
  const batch_header* h = reinterpret_cast<const batch_header*>(table);
  const batch_item_dev* items =
      reinterpret_cast<const batch_item_dev*>(table + h->items_off);
  batch_class_loop(table, kSpmvClasses + var, &s_next, [&](int item, int sub) {
    const batch_item_dev& it = items[item];
    spmm_tile<IPT, CPL, V>(it.offsets, it.indices, static_cast<const V*>(it.values), it.rows,
                           it.nnz, it.k, it.x, it.ldx, it.y, it.ldy, it.g, sub, it.coords,
                           it.sx, sm);
  });
}

// Row-block streams of the plan items: piece sub of an item is stream sub
// of its plan (padding pieces past nb * S do nothing).  The LDS and the
// stream code are the plan's own launch's.
__global__ __launch_bounds__(kBlkTpb) void batch_blocks_kernel(const char* __restrict__ table) {
  __shared__ blk_smem sm;
  __shared__ long long s_next;
  const batch_header* h = reinterpret_cast<const batch_header*>(table);
  const batch_item_dev* items =
      reinterpret_cast<const batch_item_dev*>(table + h->items_off);
  const blk_exec_args* blk = reinterpret_cast<const blk_exec_args*>(table + h->blk_off);
  batch_class_loop(table, kBlkClass, &s_next, [&](int item, int sub) {
    const batch_item_dev& it = items[item];
    const blk_exec_args& a = blk[it.blk];
    if (sub >= a.nb * a.S) return;
    const bool ahead = blk_two_ahead(a.S);
    if (it.vtype == LOOPS_VALUE_BF16) {
      if (ahead) blk_stream<bf16_t, true>(a, it.x, sub, sm);
      else blk_stream<bf16_t, false>(a, it.x, sub, sm);
    } else if (it.vtype == LOOPS_VALUE_F16) {
      if (ahead) blk_stream<f16_t, true>(a, it.x, sub, sm);
      else blk_stream<f16_t, false>(a, it.x, sub, sm);
    } else {
      if (ahead) blk_stream<float, true>(a, it.x, sub, sm);
      else blk_stream<float, false>(a, it.x, sub, sm);
    }
  });
}

// The plan items' combine: workgroup p takes 256 rows of the item found in
// the BP_BROWS prefix counts and adds each row's segment partials in the
// plan's order.
__global__ __launch_bounds__(256) void batch_blocks_combine_kernel(const char* __restrict__ table) {
  const batch_header* h = reinterpret_cast<const batch_header*>(table);
  const batch_item_dev* items =
      reinterpret_cast<const batch_item_dev*>(table + h->items_off);
  const blk_exec_args* blk = reinterpret_cast<const blk_exec_args*>(table + h->blk_off);
  const long long* P = batch_prefix_of(table, h, BP_BROWS);
  const long long p = blockIdx.x;
  const int item = batch_wave_find_item(P, h->n_items, p);
  const batch_item_dev& it = items[item];
  const blk_exec_args& a = blk[it.blk];
  const long long r = (p - P[item]) * 256 + threadIdx.x;
  if (r < a.rows) it.y[r] = blk_row_sum(a.partial, a.part_stride, a.S, a.vrow, (int)r);
}

// Merge-tile fix-up across the batch: global slot s belongs to the item
// found in the BP_TILE prefix counts.
__global__ __launch_bounds__(256) void batch_spmv_fixup_kernel(const char* __restrict__ table) {
  const batch_header* h = reinterpret_cast<const batch_header*>(table);
  const batch_item_dev* items =
      reinterpret_cast<const batch_item_dev*>(table + h->items_off);
  const long long* P = batch_prefix_of(table, h, BP_TILE);
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int item = 0;
  long long t = 0, t1 = -1;
  int r = -1;
  if (s < h->n_tiles) {
    item = batch_find_item(P, h->n_items, s);
    t = s - P[item];
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

// SpMM tile fix-up across the batch: kBatchFixLanes lanes per global tile
// slot probe its tile; a row with one tile of carries is finished by them
// (columns split between the lanes), the others by the wavefront, one after
// another.
constexpr int kBatchFixLanes = 4;
__global__ __launch_bounds__(256) void batch_spmm_fixup_kernel(const char* __restrict__ table) {
  const batch_header* h = reinterpret_cast<const batch_header*>(table);
  const batch_item_dev* items =
      reinterpret_cast<const batch_item_dev*>(table + h->items_off);
  const long long gl = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long s = gl / kBatchFixLanes;
  const int sub = (int)(gl % kBatchFixLanes);
  int item = 0, r = -1;
  long long t = 0, t_end = 0;
  if (s < h->n_units) {
    const long long* P = batch_prefix_of(table, h, BP_UNIT);
    item = batch_find_item(P, h->n_items, s);
    const batch_item_dev& it = items[item];
    t = s - P[item];
    r = spmm_fixup_probe(it.offsets, it.rows, it.ti, t, it.sx, &t_end);
  }
  const bool alone = r >= 0 && t_end == t + 1;
  if (alone)
    spmm_fixup_short_part(items[item].k, t, t_end, r, items[item].sx, items[item].y,
                          items[item].ldy, sub, kBatchFixLanes);
  unsigned long long todo = __ballot(r >= 0 && !alone && sub == 0);
  while (todo) {
    const int src = __ffsll((long long)todo) - 1;
    todo &= todo - 1;
    const batch_item_dev& it = items[__shfl(item, src, 64)];
    spmm_fixup_row(it.k, __shfl(t, src, 64), __shfl(t_end, src, 64), __shfl(r, src, 64),
                   it.sx, it.y, it.ldy);
  }
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
// SpMM lanes per unit and columns per lane: 4 columns per lane (16-byte B
// loads) when k >= kSpmmVec4MinK and B rows are 16-byte aligned, in ranges
// of the fewest lanes (at least kSpmmMinG) spanning k; else 1.
struct spmm_pick {
  int g, cpl;
};
inline spmm_pick spmm_shape_for(const loops_batch_item_t& b) {
  // the rule of loops_spmm (kSpmmVec4MinK): B's alignment only
  const bool vec4 = b.k >= kSpmmVec4MinK && b.ldx % 4 == 0 &&
                    reinterpret_cast<uintptr_t>(b.x) % 16 == 0;
  if (vec4) {
    int g = kSpmmMinG;
    while (g < 64 && g * 4 < b.k) g *= 2;
    return {g, 4};
  }
  const int g = next_pow2_le64(b.k);
  return {g < kSpmmMinG ? kSpmmMinG : g, 1};
}
}  // namespace

namespace {
// What one item contributes to a batch.
struct item_counts {
  int kind = BK_MERGE, ipt = 0, g = 0, cpl = 0, ti = 0, cls = 0;
  long long spmv_work = 0, spmm_work = 0, tiles = 0, coords = 0, units = 0;
  long long blk_work = 0, brows = 0;
  size_t spmm_ws = 0;
};

// A plan item: A's rows and cols must be the plan's; its pointers are not
// used.
loops_status_t count_plan_item(const loops_batch_item_t& b, item_counts* c) {
  const loops_spmv_plan* pl = b.plan;
  if (pl->st.layout != LOOPS_PLAN_LAYOUT_BLOCKS) return LOOPS_ERROR_NOT_SUPPORTED;
  if (b.k != 1 || b.A.rows != pl->rows || b.A.cols != pl->cols) return LOOPS_ERROR_INVALID_VALUE;
  const blk_plan_state& st = pl->st.blk;
  if ((st.rows > 0 && !b.y) || (st.nnz > 0 && !b.x)) return LOOPS_ERROR_INVALID_VALUE;
  *c = item_counts{};
  c->kind = BK_BLOCKS;
  c->cls = kBlkClass;
  if (st.rows == 0) return LOOPS_SUCCESS;  // the plan holds nothing to run
  const long long n = (long long)st.nb * st.S;
  // one segment: every XCD reads all of x, so nothing to align
  const long long pad = st.S > 1 ? kBlkPiecePad : 1;
  c->blk_work = (n + pad - 1) / pad * pad;
  c->brows = st.direct ? 0 : ((long long)st.rows + 255) / 256;  // direct: the streams write y
  return LOOPS_SUCCESS;
}

loops_status_t count_item(const loops_batch_item_t& b, item_counts* c) {
  if (b.plan) return count_plan_item(b, c);
  const loops_csr_view_t& A = b.A;
  if (A.rows < 0 || A.cols < 0 || A.nnzs < 0 || b.k < 1 || b.k > (1 << 20))
    return LOOPS_ERROR_INVALID_VALUE;
  if (A.rows + A.nnzs > LOOPS_MAX_ROWS_PLUS_NNZS || A.cols >= (int64_t)INT32_MAX)
    return LOOPS_ERROR_OVERFLOW;
  if (A.value_type != LOOPS_VALUE_F32 && A.value_type != LOOPS_VALUE_BF16 &&
      A.value_type != LOOPS_VALUE_F16)
    return LOOPS_ERROR_NOT_SUPPORTED;
  if ((A.rows > 0 && (!A.offsets || !b.y)) || (A.nnzs > 0 && (!A.indices || !A.values || !b.x)))
    return LOOPS_ERROR_INVALID_VALUE;
  const uintptr_t va = A.value_type == LOOPS_VALUE_F32 ? 16 : 8;
  if (reinterpret_cast<uintptr_t>(A.indices) % 16 || reinterpret_cast<uintptr_t>(A.values) % va)
    return LOOPS_ERROR_NOT_SUPPORTED;  // batch pieces use 16-byte loads
  if (b.k > 1 && (b.ldx < b.k || b.ldy < b.k)) return LOOPS_ERROR_INVALID_VALUE;
  const long long total = A.rows + A.nnzs;
  *c = item_counts{};
  if (b.k == 1 && b.schedule == LOOPS_SCHEDULE_THREAD_MAPPED) {
    c->kind = BK_ROWS;
    c->spmv_work = (A.rows + kBatchTpb - 1) / kBatchTpb;
  } else if (b.k == 1 && b.schedule == LOOPS_SCHEDULE_GROUP_MAPPED) {
    c->kind = BK_GROUP;
    c->g = group_lanes(b);
    if (!valid_group(c->g)) return LOOPS_ERROR_NOT_SUPPORTED;
    const long long per = kBatchTpb / c->g;
    c->spmv_work = (A.rows + per - 1) / per;
  } else if (b.k == 1) {
    if (b.schedule != LOOPS_SCHEDULE_MERGE_PATH_FLAT) return LOOPS_ERROR_NOT_SUPPORTED;
    c->kind = BK_MERGE;
    c->ipt = b.items_per_thread ? b.items_per_thread : 8;
    if (c->ipt != 4 && c->ipt != 8 && c->ipt != 16) return LOOPS_ERROR_NOT_SUPPORTED;
    c->ti = kBatchTpb * c->ipt;
    c->cls = spmv_class(c->ipt);
    const long long T = (total + c->ti - 1) / c->ti;
    c->spmv_work = T;
    c->tiles = T;
    c->coords = T + 1;
  } else {
    if (b.schedule != LOOPS_SCHEDULE_MERGE_PATH_FLAT) return LOOPS_ERROR_NOT_SUPPORTED;
    c->kind = BK_SPMM;
    const spmm_pick sp = spmm_shape_for(b);
    // 0: the per-call default (spmm_default_ipt), so an item matches
    // loops_spmm with the same parameters bit for bit
    c->ipt = b.items_per_thread ? b.items_per_thread : spmm_default_ipt(sp.g);
    if (!spmm_ipt_ok(c->ipt, sp.g)) return LOOPS_ERROR_NOT_SUPPORTED;
    c->g = sp.g;
    c->cpl = sp.cpl;
    c->ti = (int)spmm_tile_items(c->g, c->ipt);
    const long long U = (total + c->ti - 1) / c->ti;
    c->spmm_work = U;
    c->units = U;
    c->coords = U + 1;
    c->spmm_ws = U > 0 ? a256(spmm_ws_bytes_impl(U, (int)b.k)) : 0;
    c->cls = kSpmvClasses + spmm_variant(A.value_type, c->cpl, c->ipt);
  }
  return LOOPS_SUCCESS;
}
}  // namespace

namespace {
loops_status_t batch_layout_items(const loops_batch_item_t* items, int64_t n,
                                  loops_batch_info_t* info, batch_item_dev* out_table,
                                  char* dev_ws);
}  // namespace

// Sizes of the table and of the device workspace for these items; with
// out_table, also fills the table (host memory) pointing into dev_ws.  The
// table holds the header, one descriptor per item and BP_COUNT prefix
// arrays of n + 1 counts: O(n) bytes whatever the sizes of the items.
// SpMM items of more than kSpmmChunk columns become one item per column
// chunk (spmm_chunk_width), as loops_spmm launches them (each item
// validated whole first).
loops_status_t batch_layout(const loops_batch_item_t* items, int64_t n,
                            loops_batch_info_t* info, batch_item_dev* out_table,
                            char* dev_ws) {
  long long pieces = 0;
  for (int64_t i = 0; i < n; ++i) {
    item_counts c;
    const loops_status_t st = count_item(items[i], &c);
    if (st != LOOPS_SUCCESS) return st;
    if (items[i].k <= kSpmmChunk) {
      ++pieces;
      continue;
    }
    for (long long j = 0; j < items[i].k; j += spmm_chunk_width(items[i].k, j)) ++pieces;
  }
  if (pieces == n) return batch_layout_items(items, n, info, out_table, dev_ws);
  if (pieces >= (long long)INT32_MAX) return LOOPS_ERROR_OVERFLOW;
  auto part_k = [](long long k, long long j) { return (int64_t)spmm_chunk_width(k, j); };
  std::vector<loops_batch_item_t> split;
  split.reserve((size_t)pieces);
  for (int64_t i = 0; i < n; ++i) {
    const loops_batch_item_t& b = items[i];
    if (b.k <= kSpmmChunk) {
      split.push_back(b);
      continue;
    }
    for (int64_t j = 0; j < b.k; j += part_k(b.k, j)) {
      loops_batch_item_t part = b;
      part.k = part_k(b.k, j);
      part.x = b.x + j;
      part.y = b.y + j;
      split.push_back(part);
    }
  }
  return batch_layout_items(split.data(), (int64_t)split.size(), info, out_table, dev_ws);
}

namespace {
loops_status_t batch_layout_items(const loops_batch_item_t* items, int64_t n,
                                  loops_batch_info_t* info, batch_item_dev* out_table,
                                  char* dev_ws) {
  long long n_spmv_work = 0, n_spmm_work = 0, n_tiles = 0, n_coords = 0, n_units = 0;
  long long n_blk_work = 0, n_brows = 0, n_blk = 0;
  int max_ipt = 0;
  size_t spmm_ws_bytes = 0;
  long long cls_count[kClasses] = {};
  int64_t var_mask = 0, spmv_mask = 0;
  item_counts c;
  for (int64_t i = 0; i < n; ++i) {
    const loops_status_t st = count_item(items[i], &c);
    if (st != LOOPS_SUCCESS) return st;
    ++cls_count[c.cls];
    if (c.kind == BK_SPMM && c.spmm_work > 0)
      var_mask |= (int64_t)1 << (c.cls - kSpmvClasses);
    if (c.kind != BK_SPMM && c.spmv_work > 0) spmv_mask |= (int64_t)1 << c.cls;
    n_spmv_work += c.spmv_work;
    n_spmm_work += c.spmm_work;
    n_tiles += c.tiles;
    n_coords += c.coords;
    n_units += c.units;
    n_blk_work += c.blk_work;
    n_brows += c.brows;
    if (c.kind == BK_BLOCKS) ++n_blk;
    spmm_ws_bytes += c.spmm_ws;
    if (c.kind == BK_MERGE && c.ipt > max_ipt) max_ipt = c.ipt;
  }
  if (n_blk > 1) {  // the streams of one plan share its partial sums
    std::vector<const loops_spmv_plan*> plans;
    plans.reserve((size_t)n_blk);
    for (int64_t i = 0; i < n; ++i)
      if (items[i].plan) plans.push_back(items[i].plan);
    std::sort(plans.begin(), plans.end());
    if (std::adjacent_find(plans.begin(), plans.end()) != plans.end())
      return LOOPS_ERROR_INVALID_VALUE;
  }
  if (n_spmv_work >= (long long)INT32_MAX || n_spmm_work >= (long long)INT32_MAX ||
      n_tiles >= INT32_MAX || n_coords >= INT32_MAX || n_units >= INT32_MAX ||
      n_blk_work >= (long long)INT32_MAX || n_brows >= (long long)INT32_MAX || n >= INT32_MAX)
    return LOOPS_ERROR_OVERFLOW;
  const size_t tile_ws = n_tiles > 0 ? a256(fixup_ws_bytes(n_tiles)) : 0;
  const size_t coord_ws = a256((size_t)n_coords * sizeof(loops_coord_t));
  batch_header hd{};
  hd.n_items = n;
  hd.n_spmv_work = n_spmv_work;
  hd.n_spmm_work = n_spmm_work;
  hd.n_tiles = n_tiles;
  hd.n_coords = n_coords;
  hd.n_units = n_units;
  hd.n_blk_work = n_blk_work;
  hd.n_brows = n_brows;
  hd.n_blk = n_blk;
  hd.class_begin[0] = 0;
  for (int q = 0; q < kClasses; ++q) hd.class_begin[q + 1] = hd.class_begin[q] + cls_count[q];
  hd.items_off = a256(sizeof(batch_header));
  hd.prefix_off = hd.items_off + a256((size_t)n * sizeof(batch_item_dev));
  hd.order_off = hd.prefix_off + a256((size_t)BP_COUNT * (n + 1) * sizeof(long long));
  hd.cprefix_off = hd.order_off + a256((size_t)n * sizeof(int));
  hd.blk_off = hd.cprefix_off + a256((size_t)(n + 1) * sizeof(long long));
  info->table_bytes = (int64_t)(hd.blk_off + a256((size_t)n_blk * sizeof(blk_exec_args)));
  info->workspace_bytes = (int64_t)(kQueueBytes + tile_ws + coord_ws + spmm_ws_bytes + 256);
  info->work_items = n_spmv_work + n_spmm_work + n_blk_work;
  info->spmv_tiles = n_tiles;
  info->spmm_units = n_units;
  info->spmm_work_items = n_spmm_work;
  info->coordinates = n_coords;
  info->max_items_per_thread = max_ipt;
  info->spmm_variant_mask = var_mask;
  info->spmv_class_mask = spmv_mask;
  info->blocks_work_items = n_blk_work;
  info->blocks_row_pieces = n_brows;
  if (!out_table) return LOOPS_SUCCESS;

  // Second pass: fill the table.
  char* table = reinterpret_cast<char*>(out_table);
  hd.queue = reinterpret_cast<unsigned long long*>(dev_ws);
  dev_ws += kQueueBytes;
  std::memcpy(table, &hd, sizeof(hd));
  batch_item_dev* its = reinterpret_cast<batch_item_dev*>(table + hd.items_off);
  long long* pre = reinterpret_cast<long long*>(table + hd.prefix_off);
  long long* P[BP_COUNT];
  for (int q = 0; q < BP_COUNT; ++q) P[q] = pre + q * (n + 1);
  int* order = reinterpret_cast<int*>(table + hd.order_off);
  long long* CP = reinterpret_cast<long long*>(table + hd.cprefix_off);
  blk_exec_args* BA = reinterpret_cast<blk_exec_args*>(table + hd.blk_off);
  long long cls_next[kClasses];
  for (int q = 0; q < kClasses; ++q) cls_next[q] = hd.class_begin[q];
  std::vector<long long> pieces((size_t)n);
  fixup_ws all_tiles = n_tiles > 0 ? carve_fixup_ws(dev_ws, n_tiles) : fixup_ws{};
  loops_coord_t* all_coords = reinterpret_cast<loops_coord_t*>(dev_ws + tile_ws);
  char* spmm_base = dev_ws + tile_ws + coord_ws;
  long long wv = 0, wm = 0, slot = 0, coord = 0, unit = 0, brow = 0;
  int nblk = 0;
  for (int64_t i = 0; i < n; ++i) {
    const loops_batch_item_t& b = items[i];
    (void)count_item(b, &c);  // validated in the first pass
    order[cls_next[c.cls]++] = (int)i;
    pieces[i] = c.kind == BK_SPMM ? c.spmm_work : c.kind == BK_BLOCKS ? c.blk_work : c.spmv_work;
    P[BP_TILE][i] = slot;
    P[BP_COORD][i] = coord;
    P[BP_UNIT][i] = unit;
    P[BP_BROWS][i] = brow;
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
    d.kind = c.kind;
    d.ipt = c.ipt;
    d.g = c.g;
    d.cpl = c.cpl;
    d.ti = c.ti;
    d.first_slot = (int)slot;
    d.first_coord = (int)coord;
    if (c.kind == BK_BLOCKS) {
      const blk_plan_state& st = b.plan->st.blk;
      d.offsets = nullptr;
      d.indices = nullptr;
      d.values = nullptr;
      d.rows = st.rows;
      d.cols = st.cols;
      d.nnz = st.nnz;
      d.vtype = st.vtype;
      d.blk = nblk;
      BA[nblk++] = blk_args_of(st, b.y);
    }
    if (c.kind == BK_MERGE) {
      d.fx.carry_row = all_tiles.carry_row + slot;
      d.fx.seg_end = all_tiles.seg_end + slot;
      d.fx.carry_val = all_tiles.carry_val + slot;
      d.fx.head_val = all_tiles.head_val + slot;
    }
    if (c.coords > 0) d.coords = all_coords + coord;
    if (c.kind == BK_SPMM) {
      d.units = (int)c.units;
      if (c.units > 0) {
        d.sx = carve_spmm_ws(spmm_base, c.units, (int)b.k);
        spmm_base += c.spmm_ws;
      }
    }
    wv += c.spmv_work;
    wm += c.spmm_work;
    slot += c.tiles;
    coord += c.coords;
    unit += c.units;
    brow += c.brows;
    its[i] = d;
  }
  // Plan items: those of several segments first (their pieces then start
  // at multiples of kBlkPiecePad), each group heaviest first (most rounds),
  // so that the longest streams start early and short ones fill the end of
  // the launch.
  {
    int* b0 = order + hd.class_begin[kBlkClass];
    int* b1 = order + hd.class_begin[kBlkClass + 1];
    std::stable_sort(b0, b1, [&](int u, int v) {
      const blk_plan_state& a = items[u].plan->st.blk;
      const blk_plan_state& b = items[v].plan->st.blk;
      if ((a.S > 1) != (b.S > 1)) return a.S > 1;
      return a.nrounds > b.nrounds;
    });
  }
  // Pieces in class order: each class's pieces are contiguous.
  long long cp = 0;
  for (long long j = 0; j < n; ++j) {
    CP[j] = cp;
    cp += pieces[order[j]];
  }
  CP[n] = cp;
  (void)wv;
  (void)wm;
  P[BP_TILE][n] = slot;
  P[BP_COORD][n] = coord;
  P[BP_UNIT][n] = unit;
  P[BP_BROWS][n] = brow;
  return LOOPS_SUCCESS;
}
}  // namespace

namespace {
using spmm_kernel_fn = void (*)(const char*);

// The SpMM kernel of each variant, indexed by spmm_variant().
template <typename V>
void spmm_kernels_of(spmm_kernel_fn* k) {
  k[0] = batch_spmm_kernel<4, 1, V>;
  k[1] = batch_spmm_kernel<8, 1, V>;
  k[2] = batch_spmm_kernel<16, 1, V>;
  k[3] = batch_spmm_kernel<32, 1, V>;
  k[4] = batch_spmm_kernel<4, 4, V>;
  k[5] = batch_spmm_kernel<8, 4, V>;
  k[6] = batch_spmm_kernel<16, 4, V>;
  k[7] = batch_spmm_kernel<32, 4, V>;
}

struct spmm_kernel_table {
  spmm_kernel_fn fn[kSpmmVariants];
  spmm_kernel_table() {
    spmm_kernels_of<float>(fn + 8 * LOOPS_VALUE_F32);
    spmm_kernels_of<bf16_t>(fn + 8 * LOOPS_VALUE_BF16);
    spmm_kernels_of<f16_t>(fn + 8 * LOOPS_VALUE_F16);
  }
};
const spmm_kernel_table g_spmm_kernels;

// Resident workgroups of each persistent kernel on a device (occupancy x CUs),
// computed once per device under a lock (the server thread and its clients
// may ask at the same time).
struct batch_grids {
  int spmv4 = 0, spmv8 = 0, spmv16 = 0, blocks = 0;
  int spmm[kSpmmVariants] = {};
};
constexpr int kMaxDevices = 64;
batch_grids g_grids[kMaxDevices];
std::mutex g_grids_mu;

hipError_t resident_grids(batch_grids* out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(g_grids_mu);
  if (dev >= 0 && dev < kMaxDevices && g_grids[dev].spmv4 > 0) {
    *out = g_grids[dev];
    return hipSuccess;
  }
  int cus = 0, a4 = 0, a8 = 0, a16 = 0, ab = 0;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a4, batch_spmv_kernel<4>, kBatchTpb, 0);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a8, batch_spmv_kernel<8>, kBatchTpb, 0);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a16, batch_spmv_kernel<16>, kBatchTpb, 0);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&ab, batch_blocks_kernel, kBlkTpb, 0);
  if (e != hipSuccess) return e;
  batch_grids g;
  for (int v = 0; v < kSpmmVariants; ++v) {
    int b = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &b, reinterpret_cast<const void*>(g_spmm_kernels.fn[v]), kBatchTpb, 0);
    if (e != hipSuccess) return e;
    g.spmm[v] = (b > 0 ? b : 1) * cus;
  }
  g.spmv4 = (a4 > 0 ? a4 : 1) * cus;
  g.spmv8 = (a8 > 0 ? a8 : 1) * cus;
  g.spmv16 = (a16 > 0 ? a16 : 1) * cus;
  g.blocks = (ab > 0 ? ab : 1) * cus;
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
  const long long spmv_work =
      info->work_items - info->spmm_work_items - info->blocks_work_items;
  if (info->work_items > 0) {
    const long long pg = info->coordinates > 0 ? (info->coordinates + 255) / 256 : 1;
    static_assert(kQueueSets * kQueues <= 256, "prologue's first block zeroes the counters");
    batch_prologue_kernel<<<(unsigned)pg, 256, 0, s>>>(t);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (info->blocks_work_items > 0) {
    long long g = grid > 0 ? grid : rg.blocks;
    if (g > info->blocks_work_items) g = info->blocks_work_items;
    batch_blocks_kernel<<<(unsigned)g, kBlkTpb, 0, s>>>(t);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  for (int q = 0; q < kSpmvClasses && spmv_work > 0; ++q) {
    if (!((info->spmv_class_mask >> q) & 1)) continue;
    long long g = grid > 0 ? grid : (q == 2 ? rg.spmv16 : q == 1 ? rg.spmv8 : rg.spmv4);
    if (g > spmv_work) g = spmv_work;
    if (q == 2) batch_spmv_kernel<16><<<(unsigned)g, kBatchTpb, 0, s>>>(t);
    else if (q == 1) batch_spmv_kernel<8><<<(unsigned)g, kBatchTpb, 0, s>>>(t);
    else batch_spmv_kernel<4><<<(unsigned)g, kBatchTpb, 0, s>>>(t);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  for (int v = 0; v < kSpmmVariants; ++v) {
    if (!((info->spmm_variant_mask >> v) & 1)) continue;
    long long g = grid > 0 ? grid : rg.spmm[v];
    if (g > info->spmm_work_items) g = info->spmm_work_items;
    g_spmm_kernels.fn[v]<<<(unsigned)g, kBatchTpb, 0, s>>>(t);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (info->spmv_tiles > 0) {
    batch_spmv_fixup_kernel<<<(unsigned)((info->spmv_tiles + 255) / 256), 256, 0, s>>>(t);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (info->spmm_units > 0) {
    batch_spmm_fixup_kernel<<<(unsigned)((info->spmm_units * kBatchFixLanes + 255) / 256), 256,
                              0, s>>>(t);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (info->blocks_row_pieces > 0) {
    batch_blocks_combine_kernel<<<(unsigned)info->blocks_row_pieces, 256, 0, s>>>(t);
    return hipGetLastError();
  }
  return hipSuccess;
}

}  // namespace loops_mi355x
