// Row-block SpMV plan: y accumulated in LDS, x read in column order.
//
// In CSR order the nonzeros a workgroup processes together span the whole
// column range, so every x read is a separate 128-byte line request to L2
// (one line per nonzero): on the 2M-row, 40M-nonzero power-law matrix that
// traffic bounds the merge-path kernel even when x sits in L2.  This plan
// reorders A once so that the x reads of a wavefront share lines:
//
//   rows    are cut into blocks of at most kBlkRows rows and about equal
//           nonzero counts (enough blocks for two waves of resident
//           workgroups, less a few), launched heaviest first; a workgroup
//           keeps the y of its block in LDS (kBlkRows floats), two
//           workgroups per CU.  A row longer than a quarter of a block's
//           share is cut into m pieces (entry i to piece i % m) that count
//           as rows of their own;
//   columns are cut into S segments of W columns (S = 4 when 4 segments of
//           2^kBlkColBits columns hold them, else a multiple of 8), and
//           workgroup b * S + s takes segment s of block b, so a segment
//           always runs on the same XCDs (workgroups go to the 8 XCDs
//           round-robin) and each XCD's L2 holds only its slice of x;
//   entries of each (block, segment) stream are bucketed by chunks of
//           kBlkChunkCols columns and the stream is cut into rounds of
//           whole chunks, at most kBlkRound entries each, so a wavefront's
//           64 consecutive entries fall on a few x lines.
//
// A round is processed in two phases.  Phase 1 reads its entries in column
// order, multiplies by x and writes each product to its slot in LDS: the
// plan stores, per entry, the column within the segment and the entry's
// position ("slot") in the round sorted by (row, column).  Phase 2 lets lane
// t sum slots 8t .. 8t + 7 (rows from the plan's per-slot row array), a
// workgroup segmented scan carries row sums across lanes, and the lane that
// holds a row's last slot adds the row's total to the LDS accumulator --
// one writer per row per round, so no atomics, and every sum has an order
// fixed by the plan: results are bitwise reproducible.  Every workgroup
// writes its block's partial y for its segment; a combine kernel adds the S
// partials of each row in segment order (and a cut row's pieces in piece
// order).
//
// Bytes per fp32 entry: 4 (value) + 4 (column | slot) + 2 (row of a slot).
// Entries are stored lane-major in slabs of 512 (a wavefront's share of a
// round), so a lane loads its 8 column | slot words with two 16-byte loads
// and its values with two (fp32) or one (16-bit), each load a contiguous
// kilobyte across the wavefront; rounds start at whole slabs, which pads
// each round by less than 512 entries.
// The loads of round r + 1 are issued while round r's x reads are in
// flight, so each round waits for memory once.
//
// Build (loops_spmv_plan_create, layout LOOPS_PLAN_LAYOUT_BLOCKS): count
// the entries of every (block, segment, chunk) bucket, an exclusive scan,
// the rounds of every stream, a scatter of every entry into its bucket and
// a sort of every round by (row, column).  A bucket that fits a round is
// scattered with an atomic cursor (its order is then fixed by the round's
// sort); a bucket of more than kBlkRound entries is split across rounds by
// position, so a workgroup of its own places it in (row, column) order
// first.  Two builds of a matrix are therefore identical, except for the
// order of duplicate (row, column) entries inside one round.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "blk_stream.hpp"
#include "device_common.hpp"
#include "loops_internal.hpp"

namespace loops_mi355x {

// ---------------------------------------------------------------------------
// Build kernels
// ---------------------------------------------------------------------------
struct blk_geom {
  int rows, cols, nnz, S, W, nch, nb;  // nch: chunks per segment; nb: row blocks
  const int* blk_start;                // first virtual row of each block (nb + 1)
  const int* vrow;                     // first virtual row of each row, or null
  // compressed plans: exclusive prefix over (segment, virtual row) pairs,
  // segment-major, of "the row has entries in the segment" (S V + 1); null
  // for dense plans
  const int* excl;
  long long V;                         // virtual rows
  int run = 1;                         // entry kernels: entries per lane (blk_entry_run)
};

// Virtual row of entry nz of row r: the row itself, or for a row cut into
// m pieces piece (nz - row start) % m.
__device__ __forceinline__ int blk_vrow_of(const blk_geom& g, const int* __restrict__ offsets,
                                           int r, long long nz) {
  if (!g.vrow) return r;
  const int v = g.vrow[r], m = g.vrow[r + 1] - v;
  return m > 1 ? v + (int)((nz - offsets[r]) % m) : v;
}

// Block holding virtual row r: the largest b with blk_start[b] <= r.
__device__ __forceinline__ int blk_block_of(const blk_geom& g, int r) {
  int lo = 0, hi = g.nb;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (g.blk_start[mid] <= r) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int blk_row_of(const int* __restrict__ offsets, int rows,
                                          long long nz) {
  int lo = 0, hi = rows;  // offsets[lo] <= nz < offsets[lo + 1]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (offsets[mid] <= nz) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Calls f(nz, r) for every nonzero nz (row r), entries spread over the grid
// in workgroup ranges of 256 x run: lane l of a workgroup takes entries
// base + 256 j + l (coalesced), and the rows are found by a search narrowed
// to the rows of the workgroup's range, continued from the lane's previous
// row (a few cached steps instead of a search over all rows per nonzero).
// blockDim.x must be 256.
template <typename F>
__device__ __forceinline__ void blk_for_entries(const int* __restrict__ offsets, int rows,
                                                long long nnz, int run, F&& f) {
  __shared__ int s_rows[2];
  const long long per = 256LL * run;
  for (long long base = (long long)blockIdx.x * per; base < nnz;
       base += (long long)gridDim.x * per) {
    const long long last = (base + per < nnz ? base + per : nnz) - 1;
    if (threadIdx.x < 2) s_rows[threadIdx.x] = blk_row_of(offsets, rows, threadIdx.x ? last : base);
    __syncthreads();
    int r = s_rows[0];
    const int hi = s_rows[1] + 1;  // offsets[hi] > last
    __syncthreads();                // s_rows is rewritten by the next range
    for (int j = 0; j < run; ++j) {
      const long long nz = base + 256LL * j + threadIdx.x;
      if (nz > last) break;
      int lo = r, h = hi;  // offsets[lo] <= nz < offsets[h]
      while (h - lo > 1) {
        const int mid = (lo + h) >> 1;
        if (offsets[mid] <= nz) lo = mid;
        else h = mid;
      }
      r = lo;
      f(nz, r);
    }
  }
}

__device__ __forceinline__ long long blk_bucket(const blk_geom& g, int b, int col) {
  const int s = col / g.W;
  const int ch = (col - s * g.W) / kBlkChunkCols;
  return ((long long)b * g.S + s) * g.nch + ch;
}

// mask[v] |= bit s for every nonzero of virtual row v in segment s (S <= 64).
__global__ __launch_bounds__(256) void blk_mask_kernel(const int* __restrict__ offsets,
                                                       const int* __restrict__ indices,
                                                       blk_geom g,
                                                       unsigned long long* __restrict__ mask) {
  blk_for_entries(offsets, g.rows, g.nnz, g.run, [&](long long nz, int row) {
    const int v = blk_vrow_of(g, offsets, row, nz);
    const unsigned long long bit = 1ull << (indices[nz] / g.W);
    // bits are only ever set: most entries of a long row find theirs set
    // already and skip the atomic
    if (!(__hip_atomic_load(&mask[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit))
      atomicOr(&mask[v], bit);
  });
}

// fl[s V + v] = bit s of mask[v].
__global__ __launch_bounds__(256) void blk_flag_kernel(const unsigned long long* __restrict__ mask,
                                                       long long V, int S, int* __restrict__ fl) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (long long)S * V;
       i += (long long)gridDim.x * blockDim.x)
    fl[i] = (int)((mask[i % V] >> (i / V)) & 1ull);
}

// rlist[excl[s V + v]] = v for every flagged pair (the stream rows, stream
// by stream), rstart[s (nb + 1) + b] = excl[s V + blk_start[b]].
__global__ __launch_bounds__(256) void blk_rlist_kernel(const int* __restrict__ fl,
                                                        const int* __restrict__ excl, long long V,
                                                        int S, int* __restrict__ rlist) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (long long)S * V;
       i += (long long)gridDim.x * blockDim.x)
    if (fl[i]) rlist[excl[i]] = (int)(i % V);
}
__global__ __launch_bounds__(256) void blk_rstart_kernel(const int* __restrict__ excl,
                                                         const int* __restrict__ blk_start,
                                                         int nb, long long V, int S,
                                                         int* __restrict__ rstart) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= S * (nb + 1)) return;
  const int s = t / (nb + 1), b = t % (nb + 1);
  rstart[t] = excl[(long long)s * V + blk_start[b]];
}

// cnt[bucket] += 1 for every nonzero.
__global__ __launch_bounds__(256) void blk_count_kernel(const int* __restrict__ offsets,
                                                        const int* __restrict__ indices,
                                                        blk_geom g, int* __restrict__ cnt) {
  blk_for_entries(offsets, g.rows, g.nnz, g.run, [&](long long nz, int row) {
    const int r = blk_vrow_of(g, offsets, row, nz);
    atomicAdd(&cnt[blk_bucket(g, blk_block_of(g, r), indices[nz])], 1);
  });
}

// Rounds of stream q: its chunks in order, packed greedily into rounds of at
// most kBlkRound entries; a chunk larger than that is cut into full rounds
// and a remainder that later chunks may join.  With off == nullptr only
// counts them; otherwise writes each round's first entry at off[0 ...].
// One wavefront per stream: the lanes load 64 chunk starts at a time
// (coalesced) and the whole wavefront runs the packing on readlane values,
// so the sequential walk waits on memory once per 64 chunks.
__device__ __forceinline__ int blk_rounds_of(const int* __restrict__ start, long long q,
                                             int nch, int* __restrict__ off) {
  const int lane = threadIdx.x & 63;
  const long long k0 = q * nch;
  int n = 0, cur = 0, cur_start = 0;
  for (int c0 = 0; c0 < nch; c0 += 64) {
    const int c = c0 + lane;
    int a = 0, size = 0;
    if (c < nch) {
      a = start[k0 + c];
      size = start[k0 + c + 1] - a;
    }
    const int m = nch - c0 < 64 ? nch - c0 : 64;
    for (int i = 0; i < m; ++i) {
      const int ai = __builtin_amdgcn_readlane(a, i);
      const int si = __builtin_amdgcn_readlane(size, i);
      if (si == 0) continue;
      if (cur + si <= kBlkRound) {
        if (cur == 0) cur_start = ai;
        cur += si;
        continue;
      }
      if (cur > 0) {
        if (off && lane == 0) off[n] = cur_start;
        ++n;
      }
      int done = 0;
      while (si - done > kBlkRound) {
        if (off && lane == 0) off[n] = ai + done;
        ++n;
        done += kBlkRound;
      }
      cur = si - done;
      cur_start = ai + done;
    }
  }
  if (cur > 0) {
    if (off && lane == 0) off[n] = cur_start;
    ++n;
  }
  return n;
}

// One wavefront per stream (blockDim.x a multiple of 64).
__global__ __launch_bounds__(256) void blk_round_count_kernel(const int* __restrict__ start,
                                                              int nstreams, int nch,
                                                              int* __restrict__ nrounds) {
  const long long q = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (q >= nstreams) return;
  const int n = blk_rounds_of(start, q, nch, nullptr);
  if ((threadIdx.x & 63) == 0) nrounds[q] = n;
}

__global__ __launch_bounds__(256) void blk_round_write_kernel(const int* __restrict__ start,
                                                              int nstreams, int nch,
                                                              const int* __restrict__ first,
                                                              int nnz,
                                                              int* __restrict__ round_off) {
  const long long q = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (q == 0 && (threadIdx.x & 63) == 0) round_off[first[nstreams]] = nnz;
  if (q >= nstreams) return;
  blk_rounds_of(start, q, nch, round_off + first[q]);
}

// Every nonzero of a bucket of at most kBlkRound entries to the next free
// slot of its bucket (cursor = bucket starts): column within the segment to
// packed, row within the block to trow.  Such a bucket lies in one round,
// whose sort orders it by (row, column).  Larger buckets are split across
// rounds by position, so their order must not depend on scheduling:
// blk_scatter_big_kernel places them.
template <typename V>
__global__ __launch_bounds__(256) void blk_scatter_kernel(
    const int* __restrict__ offsets, const int* __restrict__ indices,
    const V* __restrict__ values, blk_geom g, const int* __restrict__ cnt,
    int* __restrict__ cursor, uint32_t* __restrict__ packed, uint16_t* __restrict__ trow,
    V* __restrict__ vals) {
  blk_for_entries(offsets, g.rows, g.nnz, g.run, [&](long long nz, int row) {
    const int r = blk_vrow_of(g, offsets, row, nz);
    const int b = blk_block_of(g, r);
    const int col = indices[nz];
    const int sg = col / g.W;
    const long long bucket = blk_bucket(g, b, col);
    if (cnt[bucket] > kBlkRound) return;
    const int pos = atomicAdd(&cursor[bucket], 1);
    packed[pos] = (uint32_t)(col - sg * g.W);
    trow[pos] = (uint16_t)(g.excl ? g.excl[(long long)sg * g.V + r] -
                                        g.excl[(long long)sg * g.V + g.blk_start[b]]
                                  : r - g.blk_start[b]);
    vals[pos] = values[nz];
  });
}

// big[...] = the buckets of more than kBlkRound entries (in any order; each
// is placed on its own), *nbig their number.
__global__ __launch_bounds__(256) void blk_list_big_kernel(const int* __restrict__ cnt,
                                                           long long nbuckets,
                                                           int* __restrict__ big,
                                                           int* __restrict__ nbig) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nbuckets;
       i += (long long)gridDim.x * blockDim.x)
    if (cnt[i] > kBlkRound) big[atomicAdd(nbig, 1)] = (int)i;
}

// One workgroup per bucket of more than kBlkRound entries: its entries in
// (virtual row, column) order from bucket start on, found row by row (each
// lane takes one virtual row of the block, finds its run in the chunk by
// binary search, and a workgroup scan of the run lengths gives the
// positions) -- the order the round sort would give, fixed before the
// bucket is cut into rounds by position.  Duplicate (row, column) entries
// keep CSR order.
constexpr int kBlkBigTpb = 1024;
template <typename V>
__global__ __launch_bounds__(kBlkBigTpb) void blk_scatter_big_kernel(
    const int* __restrict__ offsets, const int* __restrict__ indices,
    const V* __restrict__ values, blk_geom g, const int* __restrict__ big,
    const int* __restrict__ cnt, const int* __restrict__ bucket_start,
    uint32_t* __restrict__ packed, uint16_t* __restrict__ trow, V* __restrict__ vals) {
  __shared__ int s_wave[kBlkBigTpb / 64];
  const long long id = big[blockIdx.x];
  const int ch = (int)(id % g.nch);
  const int sg = (int)((id / g.nch) % g.S);
  const int b = (int)(id / ((long long)g.nch * g.S));
  const int v0 = g.blk_start[b], v1 = g.blk_start[b + 1];
  const int seg0 = sg * g.W;
  const int col_lo = seg0 + ch * kBlkChunkCols;
  int col_hi = col_lo + kBlkChunkCols;
  if (col_hi > seg0 + g.W) col_hi = seg0 + g.W;
  if (col_hi > g.cols) col_hi = g.cols;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  long long pos = bucket_start[id];
  const long long end = pos + cnt[id];  // the runs found add up to cnt[id]
  for (int vb = v0; vb < v1; vb += kBlkBigTpb) {
    const int v = vb + (int)threadIdx.x;
    int n = 0, first = 0, m = 1, o0 = 0, tr = 0;
    if (v < v1) {
      int r = v, p = 0;
      if (g.vrow) {  // the row holding virtual row v: largest r with vrow[r] <= v
        int lo = 0, hi = g.rows;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (g.vrow[mid] <= v) lo = mid;
          else hi = mid;
        }
        r = lo;
        m = g.vrow[r + 1] - g.vrow[r];
        p = v - g.vrow[r];
      }
      o0 = offsets[r];
      const int L = offsets[r + 1] - o0;
      auto lower = [&](int c) {  // first row-local index with column >= c
        int lo = 0, hi = L;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (indices[o0 + mid] < c) lo = mid + 1;
          else hi = mid;
        }
        return lo;
      };
      const int lo = lower(col_lo), hi = lower(col_hi);
      // piece p of m holds the row's entries i with i % m == p
      first = lo + ((p - lo % m) % m + m) % m;
      n = first < hi ? (hi - 1 - first) / m + 1 : 0;
      if (n > 0)
        tr = g.excl ? g.excl[(long long)sg * g.V + v] - g.excl[(long long)sg * g.V + v0]
                    : v - v0;
    }
    // exclusive scan of the run lengths over the workgroup
    int inc = n;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int up = __shfl_up(inc, d, 64);
      if (lane >= d) inc += up;
    }
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    int carry = 0, total = 0;
    for (int w = 0; w < kBlkBigTpb / 64; ++w) {
      const int t = s_wave[w];
      if (w < wave) carry += t;
      total += t;
    }
    __syncthreads();  // s_wave is rewritten by the next tile
    const long long at = pos + carry + inc - n;
    for (int k = 0; k < n && at + k < end; ++k) {
      const int nz = o0 + first + k * m;
      packed[at + k] = (uint32_t)(indices[nz] - seg0);
      trow[at + k] = (uint16_t)tr;
      vals[at + k] = values[nz];
    }
    pos += total;
  }
  // Slots the runs did not reach (none when they add up to cnt[id]) hold a
  // zero entry of the bucket's first column rather than leftover words.
  for (long long i = pos + threadIdx.x; i < end; i += kBlkBigTpb) {
    packed[i] = (uint32_t)(col_lo - seg0);
    trow[i] = 0;
    vals[i] = V{};
  }
}

// One workgroup per round: bitonic sort of its entries by (row, column,
// position) in LDS, then packed[e] = column << kBlkSlotBits | slot of e (in
// place) and srow[r * kBlkRound + slot] = row.
constexpr int kBlkSortTpb = 1024;
__global__ __launch_bounds__(kBlkSortTpb) void blk_round_sort_kernel(
    const int* __restrict__ round_off, uint32_t* __restrict__ packed,
    const uint16_t* __restrict__ trow, uint16_t* __restrict__ srow) {
  __shared__ unsigned long long key[kBlkRound];
  const int r = blockIdx.x;
  const int a = round_off[r], n = round_off[r + 1] - a;
  for (int i = threadIdx.x; i < kBlkRound; i += kBlkSortTpb)
    key[i] = i < n ? ((unsigned long long)trow[a + i] << 32) |
                         ((unsigned long long)packed[a + i] << kBlkSlotBits) | (unsigned)i
                   : ~0ULL;
  __syncthreads();
  for (int k = 2; k <= kBlkRound; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < kBlkRound; i += kBlkSortTpb) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const unsigned long long u = key[i], v = key[ixj];
          const bool up = (i & k) == 0;
          if ((u > v) == up) {
            key[i] = v;
            key[ixj] = u;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int k = threadIdx.x; k < n; k += kBlkSortTpb) {
    const unsigned long long u = key[k];
    const int i = (int)(u & (kBlkRound - 1));
    const uint32_t col = (uint32_t)(u >> kBlkSlotBits) & ((1u << kBlkColBits) - 1);
    packed[a + i] = (col << kBlkSlotBits) | (uint32_t)k;
    srow[(long long)r * kBlkRound + k] = (uint16_t)(u >> 32);
  }
}

// Stores each round lane-major: within slab q (entries 512 q .. 512 q + 511
// of the round, column order; lane l of wavefront q takes entries
// 512 q + 64 j + l, j = 0..7) a lane's 4-byte words j = 0..3 sit at
// 512 q + 4 l + j and j = 4..7 at 512 q + 256 + 4 l + j - 4, so each of its
// two 16-byte loads is one contiguous kilobyte across the wavefront; 16-bit
// values sit at 512 q + 8 l + j (one 16-byte load).  Rounds start at whole
// slabs (base[r] slabs); positions past a round's end stay 0.  One workgroup
// per round.
template <typename V>
__global__ __launch_bounds__(256) void blk_lane_order_kernel(
    const int* __restrict__ round_off, const int* __restrict__ base,
    const uint32_t* __restrict__ pk, const V* __restrict__ vv, uint32_t* __restrict__ packed,
    V* __restrict__ vals) {
  const int r = blockIdx.x;
  const int a = round_off[r], n = round_off[r + 1] - a;
  const long long b = (long long)base[r] * kBlkSlab;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int q = i / kBlkSlab, t = i % kBlkSlab, j = t / 64, l = t % 64;
    const long long s0 = b + (long long)q * kBlkSlab;
    const int w = (j / 4) * (kBlkSlab / 2) + l * 4 + j % 4;
    packed[s0 + w] = pk[a + i];
    vals[s0 + (std::is_same<V, float>::value ? w : l * kBlkIpt + j)] = vv[a + i];
  }
}

// ---------------------------------------------------------------------------
// Execute kernels (the stream and combine code is in blk_stream.hpp)
// ---------------------------------------------------------------------------

// Workgroup b * S + s runs segment s of block blk_order[b].
template <typename V, bool kAhead>
__global__ __launch_bounds__(kBlkTpb) void spmv_blocks_kernel(blk_exec_args a,
                                                              const float* __restrict__ x) {
  __shared__ blk_smem sm;
  blk_stream<V, kAhead>(a, x, (int)blockIdx.x, sm);
}

__global__ __launch_bounds__(256) void spmv_blocks_combine_kernel(
    const float* __restrict__ partial, long long part_stride, int S, int rows,
    const int* __restrict__ vrow, float* __restrict__ y) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x)
    y[r] = blk_row_sum(partial, part_stride, S, vrow, r);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {
inline size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

template <typename V>
hipError_t blk_scatter(const csr_args& A, const blk_geom& g, const int* cnt, int* cursor,
                       const int* big, int nbig, uint16_t* trow, uint32_t* pk, void* vv,
                       unsigned grid, hipStream_t s) {
  blk_scatter_kernel<V><<<grid, 256, 0, s>>>(A.offsets, A.indices,
                                             static_cast<const V*>(A.values), g, cnt, cursor,
                                             pk, trow, static_cast<V*>(vv));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || nbig == 0) return e;
  // the bucket starts of the big buckets: their cursors, untouched above
  blk_scatter_big_kernel<V><<<(unsigned)nbig, kBlkBigTpb, 0, s>>>(
      A.offsets, A.indices, static_cast<const V*>(A.values), g, big, cnt, cursor, pk, trow,
      static_cast<V*>(vv));
  return hipGetLastError();
}

template <typename V>
hipError_t blk_execute(const blk_plan_state& st, const float* x, float* y, hipStream_t s) {
  const unsigned grid = (unsigned)((long long)st.nb * st.S);
  if (blk_two_ahead(st.S))
    spmv_blocks_kernel<V, true><<<grid, kBlkTpb, 0, s>>>(blk_args_of(st, y), x);
  else
    spmv_blocks_kernel<V, false><<<grid, kBlkTpb, 0, s>>>(blk_args_of(st, y), x);
  return hipGetLastError();
}
}  // namespace

// Segments: 4 while 4 segments of 2^kBlkColBits columns hold the columns
// (each on two XCDs), else a multiple of 8 (one XCD each).
int blk_segments(long long cols) {
  const long long per = 1LL << kBlkColBits;
  if (cols <= 4 * per) return 4;
  return (int)(8 * ((cols + 8 * per - 1) / (8 * per)));
}

hipError_t blk_plan_build(const csr_args& A, blk_plan_state* st, bool for_batches,
                          hipStream_t s) {
  blk_plan_state& P = *st;
  P.rows = A.rows;
  P.cols = A.cols;
  P.nnz = A.nnz;
  P.vtype = A.vtype;
  P.S = for_batches && A.cols <= (1LL << kBlkColBits) ? 1 : blk_segments(A.cols);
  P.W = (int)((A.cols + P.S - 1) / P.S);
  if (P.W < 1) P.W = 1;
  if (A.rows == 0) return hipSuccess;  // nothing to store; execution only records ev1
  // Row blocks: about `target` nonzeros each, a row closing a block when
  // that overshoots the target less than stopping short would, and no
  // stream holding more than kBlkRows rows.  Compressed plans (at most 64
  // segments, and rows with entries in at most 3/4 of the (segment, row)
  // pairs) count a row only in the segments where it has entries; dense
  // plans count every row in every segment.  The cut aims at one wave of
  // big blocks (resident workgroups, two per CU, less 1/32 so every XCD
  // keeps a few slots spare; big blocks put more of a stream's entries on
  // each x line) when the row capacity allows; a one-wave cut more uneven
  // than 1.2 x is redone for two waves.  The target grows until the cut yields no more
  // blocks than wanted.  A row longer than an eighth of the one-wave share
  // (and than a round) is cut into pieces, entry i of m to piece i % m, so
  // that every piece spans the columns as the row does.  Plans for batch
  // items aim at the fewest blocks the row capacity allows instead (the
  // batch fills the GPU), with shares and the piece length to match, and
  // are never compressed when they have one segment.  A batch's persistent
  // workgroups take streams one after another, so a stream is kept to about
  // kBlkBatchRounds rounds: on 512 power-law matrices of 1e3-2e5 rows
  // (Zipf(1.8) row lengths, 253M nonzeros, plans for the 367 of 2^17 or more
  // nonzeros) a batch pass took 1.42 ms with blocks of the full row
  // capacity (streams of up to 54 rounds: the last ones ran alone), 1.21
  // with 32 rounds, 1.01 with 16, 0.94 with 8, 0.96 with 4; merge-path
  // tiles alone 1.28 ms.
  std::vector<int> off_h((size_t)A.rows + 1, 0), starts(1, 0);
  hipError_t e = hipMemcpyAsync(off_h.data(), A.offsets, ((size_t)A.rows + 1) * sizeof(int),
                                hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  int dev = 0, cus = 0;
  if (e == hipSuccess) e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  const long long wgs = 2LL * cus * 31 / 32;
  // batch items: blocks of about kBlkBatchRounds rounds per stream, or
  // fewer rows when the LDS row capacity binds
  const long long per_block = (long long)kBlkBatchRounds * kBlkRound * P.S;
  const long long fewest = std::max(((long long)A.rows + kBlkRows - 1) / kBlkRows,
                                    ((long long)A.nnz + per_block - 1) / per_block);
  const long long want1 = for_batches ? fewest  // the fewest blocks
                                      : (wgs / P.S > 0 ? wgs / P.S : 1);  // blocks of one wave
  const long long share1 = A.nnz / want1 > 0 ? A.nnz / want1 : 1;
  const long long cap = share1 / 8 > kBlkRound ? share1 / 8 : kBlkRound;
  std::vector<int> vrow_h;
  long long V = A.rows;
  for (int r = 0; r < A.rows; ++r) {
    const long long L = off_h[r + 1] - off_h[r];
    if (L > cap) {
      V += (L + cap - 1) / cap - 1;
      ++P.nsplit;
    }
  }
  if (V >= (long long)INT32_MAX) return hipErrorInvalidValue;
  P.vrows = (int)V;
  if (P.nsplit > 0) {
    vrow_h.resize((size_t)A.rows + 1);
    int v = 0;
    for (int r = 0; r < A.rows; ++r) {
      const long long L = off_h[r + 1] - off_h[r];
      vrow_h[r] = v;
      v += L > cap ? (int)((L + cap - 1) / cap) : 1;
    }
    vrow_h[A.rows] = v;
  }
  const bool masks = P.S > 1 && P.S <= 64 && (long long)P.S * V <= (1LL << 28);

  // Early scratch, kept to the end: first virtual row of each row, segment
  // masks of the virtual rows (at most 64 segments).
  const size_t b_vrow = al256(vrow_h.size() * sizeof(int));
  const size_t b_mask = masks ? al256((size_t)V * sizeof(unsigned long long)) : 0;
  char* sc0 = nullptr;
  if ((e = hipMalloc(&sc0, b_vrow + b_mask + 256)) != hipSuccess) return e;
  int* vrow = P.nsplit > 0 ? reinterpret_cast<int*>(sc0) : nullptr;
  unsigned long long* mask = masks ? reinterpret_cast<unsigned long long*>(sc0 + b_vrow) : nullptr;
  // Entry kernels (blk_for_entries): up to 16 entries per lane, as long as
  // that leaves 2,048 workgroups or more (small matrices take one per lane,
  // so the whole GPU searches in parallel).
  const int run = (int)std::min<long long>(16, std::max<long long>(1, (long long)A.nnz / (256LL * 2048)));
  long long grid = ((long long)A.nnz + 256LL * run - 1) / (256LL * run);
  if (grid > 65536) grid = 65536;
  if (grid < 1) grid = 1;
  // filled by the copy below: not value-initialised first
  std::unique_ptr<unsigned long long[]> mask_h;
  if (vrow)
    e = hipMemcpyAsync(vrow, vrow_h.data(), vrow_h.size() * sizeof(int), hipMemcpyHostToDevice,
                       s);
  if (masks && e == hipSuccess) {
    e = hipMemsetAsync(mask, 0, (size_t)V * sizeof(unsigned long long), s);
    blk_geom g0{A.rows, A.cols, A.nnz, P.S, P.W, 0, 0, nullptr, vrow, nullptr, V};
    g0.run = run;
    if (e == hipSuccess && A.nnz > 0) {
      blk_mask_kernel<<<(unsigned)grid, 256, 0, s>>>(A.offsets, A.indices, g0, mask);
      e = hipGetLastError();
    }
    mask_h.reset(new unsigned long long[(size_t)V]);
    if (e == hipSuccess)
      e = hipMemcpyAsync(mask_h.get(), mask, (size_t)V * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  if (e != hipSuccess) {
    (void)hipFree(sc0);
    return e;
  }

  // Segment masks of up to 4 segments as 16-bit lanes: spread[m] has a 1 in
  // lane s for every bit s of m.
  unsigned long long spread[16];
  for (int m = 0; m < 16; ++m) {
    spread[m] = 0;
    for (int sg = 0; sg < 4; ++sg)
      if ((m >> sg) & 1) spread[m] |= 1ull << (16 * sg);
  }
  static_assert(kBlkRows < (1 << 15), "row counts fit the 15 low bits of a 16-bit lane");
  // Compressed only where it relaxes the row capacity: when rows have
  // entries in at most 3/4 of the (segment, row) pairs.
  std::vector<long long> rows_in_seg((size_t)P.S, 0);
  long long flagged = 0;
  if (masks && P.S <= 4) {
    // 16-bit lanes of row counts per segment, one add per row, emptied into
    // 64-bit totals before a lane can overflow
    unsigned long long lanes = 0;
    auto flush = [&] {
      for (int sg = 0; sg < P.S; ++sg) rows_in_seg[sg] += (long long)((lanes >> (16 * sg)) & 0xffffu);
      lanes = 0;
    };
    for (long long v = 0; v < V; ++v) {
      lanes += spread[mask_h[v] & 15u];
      if ((v & 0x7fff) == 0x7fff) flush();
    }
    flush();
  } else if (masks) {
    for (long long v = 0; v < V; ++v)
      for (unsigned long long mm = mask_h[v]; mm; mm &= mm - 1) ++rows_in_seg[__builtin_ctzll(mm)];
  }
  for (long long c : rows_in_seg) flagged += c;
  const bool comp = masks && flagged * 4 <= 3LL * P.S * V;
  // Nonzeros of every virtual row (a row cut into m pieces: piece i holds
  // entries i, i + m, ...), so the cut below divides nothing.
  std::vector<int> vlen((size_t)V);
  for (long long r = 0, v = 0; r < A.rows; ++r) {
    const long long L = off_h[r + 1] - off_h[r];
    if (L <= cap) {
      vlen[v++] = (int)L;
      continue;
    }
    const long long m = (L + cap - 1) / cap;
    for (long long i = 0; i < m; ++i) vlen[v++] = (int)((L - i + m - 1) / m);
  }
  std::vector<long long> bnnz;
  // Greedy cut: a virtual row closes the block before it when the block
  // would exceed its row capacity (compressed: some segment of the row
  // already holds kBlkRows stream rows) or when stopping short of the
  // target is closer to it than overshooting.  With up to 4 segments the
  // per-segment row counts are 16-bit lanes of one word, so the capacity
  // test is a branch-free zero-lane test (the same cut as counting segment
  // by segment, 4x faster on the 2M-row matrix).
  const unsigned long long cap_lanes = spread[15] * (unsigned long long)kBlkRows;
  const bool packed4 = comp && P.S <= 4;
  std::vector<int> cs((size_t)P.S);
  auto cut = [&](long long want) {
    long long target = A.nnz / want > 0 ? A.nnz / want : 1;
    size_t last = 0, stalled = 0;
    for (int tries = 0;; ++tries) {
      starts.assign(1, 0);
      bnnz.clear();
      std::fill(cs.begin(), cs.end(), 0);
      unsigned long long lanes = 0;
      int rows_in = 0;
      long long nnz_in = 0;
      for (long long v = 0; v < V; ++v) {
        const long long L = vlen[v];
        bool full = rows_in == kBlkRows;
        if (packed4) {
          // lanes equal to kBlkRows: the high bit of each zero lane of x
          const unsigned long long x = lanes ^ cap_lanes;
          const unsigned long long z =
              ~(((x & 0x7fff7fff7fff7fffull) + 0x7fff7fff7fff7fffull) | x) & 0x8000800080008000ull;
          const unsigned at_cap = (unsigned)((z >> 15) & 1) | (unsigned)((z >> 30) & 2) |
                                  (unsigned)((z >> 45) & 4) | (unsigned)((z >> 60) & 8);
          full = ((unsigned)mask_h[v] & at_cap) != 0;
        } else if (comp) {
          full = false;
          for (unsigned long long mm = mask_h[v]; mm; mm &= mm - 1)
            full = full || cs[__builtin_ctzll(mm)] == kBlkRows;
        }
        if (rows_in > 0 &&
            (full || (nnz_in + L > target && nnz_in + L - target > target - nnz_in))) {
          starts.push_back((int)v);
          bnnz.push_back(nnz_in);
          rows_in = 0;
          nnz_in = 0;
          lanes = 0;
          std::fill(cs.begin(), cs.end(), 0);
        }
        if (packed4)
          lanes += spread[mask_h[v] & 15u];
        else if (comp)
          for (unsigned long long mm = mask_h[v]; mm; mm &= mm - 1) ++cs[__builtin_ctzll(mm)];
        ++rows_in;
        nnz_in += L;
      }
      starts.push_back((int)V);
      bnnz.push_back(nnz_in);
      if ((long long)bnnz.size() <= want || tries == 63) break;
      // the row capacity, not the target, sets the count: stop when it no
      // longer falls
      stalled = bnnz.size() < last ? 0 : stalled + 1;
      last = bnnz.size();
      if (stalled == 3) break;
      target += target / 64 + 1;
    }
  };
  {
    // blocks the row capacity needs (compressed: the segment with the most
    // rows); one wave when that fits and the cut is even, else two
    long long need = (V + kBlkRows - 1) / kBlkRows;
    if (comp) {
      need = 1;
      for (long long c : rows_in_seg) need = std::max(need, (c + kBlkRows - 1) / kBlkRows);
    }
    const long long slots = 2LL * cus / P.S;  // blocks of one full wave
    if (for_batches) {
      cut(std::max(need, fewest));
    } else {
      long long want = std::max(want1, need);
      if (want > slots) want = std::max(2 * want1, need);
      cut(want);
      long long most = 0;
      for (long long n : bnnz) most = n > most ? n : most;
      if (want <= slots && (double)most * bnnz.size() > 1.2 * (double)A.nnz)
        cut(std::max(2 * want1, need));
    }
  }
  P.direct = P.S == 1 && P.nsplit == 0 && !comp;
  P.nb = (int)starts.size() - 1;
  // Launch order: heaviest blocks first.
  std::vector<int> order((size_t)P.nb);
  for (int b = 0; b < P.nb; ++b) order[b] = b;
  std::stable_sort(order.begin(), order.end(),
                   [&](int u, int w) { return bnnz[u] > bnnz[w]; });
  const int nch = (P.W + kBlkChunkCols - 1) / kBlkChunkCols;
  const long long nstreams = (long long)P.nb * P.S;
  const long long nbuckets = nstreams * nch;
  if (nbuckets + 1 >= (long long)INT32_MAX) {
    (void)hipFree(sc0);
    return hipErrorInvalidValue;
  }
  const int vbytes = A.vtype == LOOPS_VALUE_F32 ? 4 : 2;
  P.part_stride = (V + 63) & ~63LL;

  // Build scratch: bucket counts, bucket starts (+1), rounds per stream and
  // their exclusive prefix (+1), scan scratch, rows of entries in stream
  // order, block starts, (compressed) pair flags and their exclusive prefix.
  const long long npairs = comp ? (long long)P.S * V : 0;
  long long nscan = nbuckets > nstreams ? nbuckets : nstreams;
  if (npairs > nscan) nscan = npairs;
  const size_t b_cnt = al256((size_t)nbuckets * sizeof(int));
  const size_t b_start = al256((size_t)(nbuckets + 1) * sizeof(int));
  const size_t b_nr = al256((size_t)(nstreams + 1) * sizeof(int));
  const size_t b_scan = al256(exclusive_scan_scratch_bytes(nscan));
  const size_t b_trow = al256(((size_t)A.nnz + 8) * sizeof(uint16_t));
  const size_t b_bstart = al256(((size_t)P.nb + 1) * sizeof(int));
  const size_t b_order = al256(((size_t)P.nb + 1) * sizeof(int));
  const size_t b_big = al256(((size_t)A.nnz / (kBlkRound + 1) + 2) * sizeof(int));
  const size_t b_fl = comp ? al256((size_t)npairs * sizeof(int)) : 0;
  const size_t b_excl = comp ? al256((size_t)(npairs + 1) * sizeof(int)) : 0;
  // Entries in stream order (column | slot words, values) until they are
  // stored lane-major in the plan.
  const size_t n_pad = (size_t)A.nnz + 8;
  const size_t b_work = al256(n_pad * sizeof(uint32_t));
  const size_t b_workv = al256(n_pad * vbytes);
  size_t sc_bytes = 0;
  auto carve = [&](size_t bytes) {
    const size_t at = sc_bytes;
    sc_bytes += bytes;
    return at;
  };
  const size_t c_cnt = carve(b_cnt), c_start = carve(b_start), c_nr = carve(b_nr),
               c_first = carve(b_nr), c_scan = carve(b_scan), c_trow = carve(b_trow),
               c_bstart = carve(b_bstart), c_big = carve(b_big), c_fl = carve(b_fl),
               c_excl = carve(b_excl),
               c_work = carve(b_work), c_workv = carve(b_workv);
  (void)c_workv;
  char* sc = nullptr;
  if ((e = hipMalloc(&sc, sc_bytes)) != hipSuccess) {
    (void)hipFree(sc0);
    return e;
  }
  int* cnt = reinterpret_cast<int*>(sc + c_cnt);
  int* start = reinterpret_cast<int*>(sc + c_start);
  int* nr = reinterpret_cast<int*>(sc + c_nr);
  int* first = reinterpret_cast<int*>(sc + c_first);
  void* scan_scratch = sc + c_scan;
  uint16_t* trow = reinterpret_cast<uint16_t*>(sc + c_trow);
  int* bstart = reinterpret_cast<int*>(sc + c_bstart);
  int* big = reinterpret_cast<int*>(sc + c_big);  // [0]: count, then the bucket ids
  int* fl = comp ? reinterpret_cast<int*>(sc + c_fl) : nullptr;
  int* excl = comp ? reinterpret_cast<int*>(sc + c_excl) : nullptr;
  char* work = sc + c_work;
  blk_geom g{A.rows, A.cols, A.nnz, P.S, P.W, nch, P.nb, bstart, vrow, excl, V};
  g.run = run;
  const unsigned sgrid = (unsigned)((nstreams + 3) / 4 > 0 ? (nstreams + 3) / 4 : 1);  // wave per stream
  auto fail = [&](hipError_t err) {
    (void)hipFree(sc);
    (void)hipFree(sc0);
    return err;
  };

  // Plan memory, part 1: stream -> first round (needed to count rounds);
  // compressed plans: the pair flags, their prefix and its total.
  const size_t b_sround = al256((size_t)(nstreams + 1) * sizeof(int));
  int nrounds = 0, np = 0;
  e = hipMemcpyAsync(bstart, starts.data(), starts.size() * sizeof(int), hipMemcpyHostToDevice,
                     s);
  if (comp && e == hipSuccess) {
    long long fgrid = (npairs + 255) / 256;
    if (fgrid > 65536) fgrid = 65536;
    blk_flag_kernel<<<(unsigned)fgrid, 256, 0, s>>>(mask, V, P.S, fl);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_exclusive_scan(fl, npairs, excl, scan_scratch, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(&np, excl + npairs, sizeof(int), hipMemcpyDeviceToHost, s);
  }
  if (e == hipSuccess) e = hipMemsetAsync(cnt, 0, b_cnt, s);
  if (e == hipSuccess && A.nnz > 0) {
    blk_count_kernel<<<(unsigned)grid, 256, 0, s>>>(A.offsets, A.indices, g, cnt);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = launch_exclusive_scan(cnt, nbuckets, start, scan_scratch, s);
  // buckets of more than a round, placed deterministically by the scatter
  int nbig = 0;
  if (e == hipSuccess) e = hipMemsetAsync(big, 0, sizeof(int), s);
  if (e == hipSuccess && A.nnz > kBlkRound) {
    long long lgrid = (nbuckets + 255) / 256;
    if (lgrid > 65536) lgrid = 65536;
    blk_list_big_kernel<<<(unsigned)lgrid, 256, 0, s>>>(cnt, nbuckets, big + 1, big);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&nbig, big, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) {
    blk_round_count_kernel<<<sgrid, 256, 0, s>>>(start, (int)nstreams, nch, nr);
    e = hipGetLastError();
  }
  // first round of each stream (copied into the plan below)
  if (e == hipSuccess) e = launch_exclusive_scan(nr, nstreams, first, scan_scratch, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(&nrounds, first + nstreams, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(e);
  P.nrounds = nrounds;

  // Plan memory, part 2: entries lane-major in whole slabs per round (at
  // most 511 positions of padding per round), rows per slot (kBlkRound per
  // round), rounds and their first slabs, partials, blocks.
  const size_t n_lane = (size_t)A.nnz + (size_t)nrounds * (kBlkSlab - 1) + 8;
  const size_t b_packed = al256(n_lane * sizeof(uint32_t));
  const size_t b_vals = al256(n_lane * vbytes);
  const size_t b_srow = al256(((size_t)nrounds * kBlkRound + 8) * sizeof(uint16_t));
  const size_t b_roff = al256((size_t)(nrounds + 1) * sizeof(int));
  const size_t b_part = P.direct ? 0 : al256((size_t)P.S * P.part_stride * sizeof(float));
  size_t total = 0;
  auto take = [&](size_t bytes) {
    const size_t at = total;
    total += bytes;
    return at;
  };
  const size_t o_packed = take(b_packed), o_vals = take(b_vals), o_srow = take(b_srow);
  const size_t o_sround = take(b_sround), o_roff = take(b_roff), o_rbase = take(b_roff);
  const size_t o_part = take(b_part), o_bstart = take(b_bstart), o_order = take(b_order);
  const size_t o_vrow = take(b_vrow);
  const size_t o_rstart = take(comp ? al256((size_t)P.S * (P.nb + 1) * sizeof(int)) : 0);
  const size_t o_rlist = take(comp ? al256(((size_t)np + 1) * sizeof(int)) : 0);
  if ((e = hipMalloc(&P.mem, total)) != hipSuccess) return fail(e);
  char* m = static_cast<char*>(P.mem);
  P.packed = reinterpret_cast<uint32_t*>(m + o_packed);
  P.values = m + o_vals;
  P.srow = reinterpret_cast<uint16_t*>(m + o_srow);
  P.stream_round = reinterpret_cast<int*>(m + o_sround);
  P.round_off = reinterpret_cast<int*>(m + o_roff);
  P.round_base = reinterpret_cast<int*>(m + o_rbase);
  P.partial = P.direct ? nullptr : reinterpret_cast<float*>(m + o_part);
  P.blk_start = reinterpret_cast<int*>(m + o_bstart);
  P.blk_order = reinterpret_cast<int*>(m + o_order);
  if (vrow) P.vrow = reinterpret_cast<int*>(m + o_vrow);
  if (comp) {
    P.rstart = reinterpret_cast<int*>(m + o_rstart);
    P.rlist = reinterpret_cast<int*>(m + o_rlist);
  }
  P.device_bytes = total;

  e = hipMemsetAsync(P.mem, 0, b_packed + b_vals + b_srow, s);
  // compressed plans write the partials of their stream rows only: the
  // others stay 0
  if (comp && e == hipSuccess) e = hipMemsetAsync(P.partial, 0, b_part, s);
  if (comp && e == hipSuccess) {
    const int t = P.S * (P.nb + 1);
    blk_rstart_kernel<<<(unsigned)((t + 255) / 256), 256, 0, s>>>(excl, bstart, P.nb, V, P.S,
                                                                 P.rstart);
    e = hipGetLastError();
    long long fgrid = (npairs + 255) / 256;
    if (fgrid > 65536) fgrid = 65536;
    if (e == hipSuccess) {
      blk_rlist_kernel<<<(unsigned)fgrid, 256, 0, s>>>(fl, excl, V, P.S, P.rlist);
      e = hipGetLastError();
    }
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(P.stream_round, first, (size_t)(nstreams + 1) * sizeof(int),
                       hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(P.blk_start, bstart, ((size_t)P.nb + 1) * sizeof(int),
                       hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(P.blk_order, order.data(), (size_t)P.nb * sizeof(int),
                       hipMemcpyHostToDevice, s);
  if (e == hipSuccess && vrow)
    e = hipMemcpyAsync(P.vrow, vrow, vrow_h.size() * sizeof(int), hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) {
    blk_round_write_kernel<<<sgrid, 256, 0, s>>>(start, (int)nstreams, nch, P.stream_round,
                                                 A.nnz, P.round_off);
    e = hipGetLastError();
  }
  // First slab of every round (host prefix over the rounds' slab counts).
  std::vector<int> roff_h((size_t)nrounds + 1), rbase_h((size_t)nrounds + 1);
  if (e == hipSuccess)
    e = hipMemcpyAsync(roff_h.data(), P.round_off, roff_h.size() * sizeof(int),
                       hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess) {
    long long slabs = 0;
    for (int r = 0; r < nrounds; ++r) {
      rbase_h[r] = (int)slabs;
      slabs += (roff_h[r + 1] - roff_h[r] + kBlkSlab - 1) / kBlkSlab;
    }
    rbase_h[nrounds] = (int)slabs;
    e = hipMemcpyAsync(P.round_base, rbase_h.data(), rbase_h.size() * sizeof(int),
                       hipMemcpyHostToDevice, s);
  }
  // the bucket starts become the scatter cursors
  uint32_t* pk = reinterpret_cast<uint32_t*>(work);
  void* vv = work + b_work;
  if (e == hipSuccess && A.nnz > 0) {
    switch (A.vtype) {
      case LOOPS_VALUE_F32:
        e = blk_scatter<float>(A, g, cnt, start, big + 1, nbig, trow, pk, vv, (unsigned)grid, s);
        break;
      case LOOPS_VALUE_BF16:
        e = blk_scatter<bf16_t>(A, g, cnt, start, big + 1, nbig, trow, pk, vv, (unsigned)grid, s);
        break;
      case LOOPS_VALUE_F16:
        e = blk_scatter<f16_t>(A, g, cnt, start, big + 1, nbig, trow, pk, vv, (unsigned)grid, s);
        break;
      default:  // storage types of the typed entry points never reach a plan
        e = hipErrorInvalidValue;
        break;
    }
  }
  if (e == hipSuccess && nrounds > 0) {
    blk_round_sort_kernel<<<(unsigned)nrounds, kBlkSortTpb, 0, s>>>(P.round_off, pk, trow,
                                                                     P.srow);
    e = hipGetLastError();
  }
  if (e == hipSuccess && nrounds > 0) {
    switch (A.vtype) {
      case LOOPS_VALUE_F32:
        blk_lane_order_kernel<float><<<(unsigned)nrounds, 256, 0, s>>>(
            P.round_off, P.round_base, pk, static_cast<const float*>(vv), P.packed,
            static_cast<float*>(P.values));
        break;
      case LOOPS_VALUE_BF16:
        blk_lane_order_kernel<bf16_t><<<(unsigned)nrounds, 256, 0, s>>>(
            P.round_off, P.round_base, pk, static_cast<const bf16_t*>(vv), P.packed,
            static_cast<bf16_t*>(P.values));
        break;
      case LOOPS_VALUE_F16:
        blk_lane_order_kernel<f16_t><<<(unsigned)nrounds, 256, 0, s>>>(
            P.round_off, P.round_base, pk, static_cast<const f16_t*>(vv), P.packed,
            static_cast<f16_t*>(P.values));
        break;
      default:
        e = hipErrorInvalidValue;
        break;
    }
    e = hipGetLastError();
  }
  const hipError_t e2 = hipStreamSynchronize(s);
  (void)hipFree(sc);
  (void)hipFree(sc0);
  return e != hipSuccess ? e : e2;
}

hipError_t blk_plan_execute(const blk_plan_state& st, const float* x, float* y,
                            hipEvent_t ev1, hipStream_t s) {
  hipError_t e = hipSuccess;
  if (st.rows == 0) return ev1 ? hipEventRecord(ev1, s) : hipSuccess;
  switch (st.vtype) {
    case LOOPS_VALUE_F32: e = blk_execute<float>(st, x, y, s); break;
    case LOOPS_VALUE_BF16: e = blk_execute<bf16_t>(st, x, y, s); break;
    case LOOPS_VALUE_F16: e = blk_execute<f16_t>(st, x, y, s); break;
    default: e = hipErrorInvalidValue; break;
  }
  if (e != hipSuccess) return e;
  if (ev1 && (e = hipEventRecord(ev1, s)) != hipSuccess) return e;
  if (st.direct) return hipSuccess;  // the streams wrote y
  long long grid = ((long long)st.rows + 255) / 256;
  if (grid > 8192) grid = 8192;
  spmv_blocks_combine_kernel<<<(unsigned)grid, 256, 0, s>>>(st.partial, st.part_stride, st.S,
                                                             st.rows, st.vrow, y);
  return hipGetLastError();
}

void blk_plan_release(blk_plan_state* st) {
  if (st->mem) (void)hipFree(st->mem);
  st->mem = nullptr;
}

}  // namespace loops_mi355x
