// Row-block SpMV plan: y accumulated in LDS, x read in column order.
//
// In CSR order the nonzeros a workgroup processes together span the whole
// column range, so every x read is a separate 128-byte line request to L2
// (one line per nonzero): on the 2M-row, 40M-nonzero power-law matrix that
// traffic bounds the merge-path kernel even when x sits in L2.  This plan
// reorders A once so that the x reads of a wavefront share lines:
//
//   rows    are cut into blocks of kBlkRows rows; a workgroup keeps the y
//           of its block in LDS (kBlkRows floats) and adds products into it
//           with LDS atomics;
//   columns are cut into S segments of W columns (S a multiple of 8), and
//           workgroup b * S + s takes segment s of block b, so segment s
//           runs on XCD s % 8 (workgroups go to the 8 XCDs round-robin) and
//           each XCD's L2 holds only its slice of x;
//   entries of each (block, segment) stream are bucketed by chunks of
//           kBlkChunkCols columns, so a wavefront's 64 consecutive entries
//           fall on a few x lines instead of 64.
//
// Each entry is stored as its value and one 32-bit word: column within the
// segment (kBlkColBits bits) above row within the block (kBlkRowBits bits)
// -- 8 bytes per fp32 entry, as in CSR.  Every workgroup writes its block's
// partial y for its segment; a combine kernel adds the S partials of each
// row in segment order.  LDS float atomics add in an order that depends on
// scheduling, so results are not bitwise reproducible from run to run (each
// row's error stays within (n + 1) u sum|a x|, like any summation order);
// the merge-path paths are the deterministic ones.
//
// Build (loops_spmv_plan_create, layout LOOPS_PLAN_LAYOUT_BLOCKS): count the
// entries of every (block, segment, chunk) bucket, an exclusive scan over
// the buckets, then a scatter of every entry into its bucket.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "device_common.hpp"
#include "loops_internal.hpp"

namespace loops_mi355x {

// ---------------------------------------------------------------------------
// Build kernels
// ---------------------------------------------------------------------------
struct blk_geom {
  int rows, cols, nnz, S, W, nch, nb;  // nch: chunks per segment; nb: row blocks
};

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

__device__ __forceinline__ long long blk_bucket(const blk_geom& g, int r, int col) {
  const int b = r >> kBlkRowBits;
  const int s = col / g.W;
  const int ch = (col - s * g.W) / kBlkChunkCols;
  return ((long long)b * g.S + s) * g.nch + ch;
}

// cnt[bucket] += 1 for every nonzero.
__global__ __launch_bounds__(256) void blk_count_kernel(const int* __restrict__ offsets,
                                                        const int* __restrict__ indices,
                                                        blk_geom g, int* __restrict__ cnt) {
  for (long long nz = (long long)blockIdx.x * blockDim.x + threadIdx.x; nz < g.nnz;
       nz += (long long)gridDim.x * blockDim.x) {
    const int r = blk_row_of(offsets, g.rows, nz);
    atomicAdd(&cnt[blk_bucket(g, r, indices[nz])], 1);
  }
}

// Every nonzero to the next free slot of its bucket (cursor = bucket starts).
template <typename V>
__global__ __launch_bounds__(256) void blk_scatter_kernel(
    const int* __restrict__ offsets, const int* __restrict__ indices,
    const V* __restrict__ values, blk_geom g, int* __restrict__ cursor,
    uint32_t* __restrict__ packed, V* __restrict__ vals) {
  for (long long nz = (long long)blockIdx.x * blockDim.x + threadIdx.x; nz < g.nnz;
       nz += (long long)gridDim.x * blockDim.x) {
    const int r = blk_row_of(offsets, g.rows, nz);
    const int col = indices[nz];
    const int pos = atomicAdd(&cursor[blk_bucket(g, r, col)], 1);
    const int s = col / g.W;
    packed[pos] = ((uint32_t)(col - s * g.W) << kBlkRowBits) |
                  (uint32_t)(r & (kBlkRows - 1));
    vals[pos] = values[nz];
  }
}

// stream_off[q] = start of the first bucket of stream q = (block, segment).
__global__ void blk_streams_kernel(const int* __restrict__ starts, int nstreams, int nch,
                                   int* __restrict__ stream_off) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q <= nstreams) stream_off[q] = starts[(long long)q * nch];
}

// ---------------------------------------------------------------------------
// Execute kernels
// ---------------------------------------------------------------------------
constexpr int kBlkTpb = 512;
constexpr int kBlkVec = 4;  // 16-byte loads of packed words: 4 entries
constexpr int kBlkU = 4;    // vector loads per lane per round: 16 entries in flight

template <typename V>
struct blk_vec;
template <>
struct blk_vec<float> {
  __device__ static __forceinline__ f32x4 load(const float* p) {
    return value_traits<float>::load4(p);
  }
};
template <>
struct blk_vec<bf16_t> {
  __device__ static __forceinline__ f32x4 load(const bf16_t* p) {
    return value_traits<bf16_t>::load4(p);
  }
};
template <>
struct blk_vec<f16_t> {
  __device__ static __forceinline__ f32x4 load(const f16_t* p) {
    return value_traits<f16_t>::load4(p);
  }
};

// Workgroup b * S + s: segment s of row block b.  Entries are read in
// aligned groups of four covering [e0, e1) (the groups may start up to 3
// entries before e0 and end up to 3 after e1; those lanes add nothing, and
// the arrays are padded so the reads stay in bounds).
template <typename V>
__global__ __launch_bounds__(kBlkTpb) void spmv_blocks_kernel(
    const uint32_t* __restrict__ packed, const V* __restrict__ vals,
    const int* __restrict__ stream_off, const float* __restrict__ x, int rows, int S,
    int W, float* __restrict__ partial, long long part_stride) {
  __shared__ float acc[kBlkRows];
  const int q = blockIdx.x;
  const int b = q / S, s = q - b * S;
  const int r0 = b << kBlkRowBits;
  const int nrows = rows - r0 < kBlkRows ? rows - r0 : kBlkRows;
  for (int i = threadIdx.x; i < nrows; i += kBlkTpb) acc[i] = 0.f;
  __syncthreads();
  const long long e0 = stream_off[q], e1 = stream_off[q + 1];
  const float* __restrict__ xs = x + (long long)s * W;
  const long long g0 = e0 & ~(long long)(kBlkVec - 1);
  for (long long base = g0; base < e1; base += (long long)kBlkTpb * kBlkVec * kBlkU) {
    i32x4 pw[kBlkU];
    f32x4 vv[kBlkU];
#pragma unroll
    for (int u = 0; u < kBlkU; ++u) {
      const long long e = base + ((long long)u * kBlkTpb + threadIdx.x) * kBlkVec;
      if (e < e1) {
        pw[u] = ld_idx4(reinterpret_cast<const int*>(packed + e));
        vv[u] = blk_vec<V>::load(vals + e);
      }
    }
    float xv[kBlkU][kBlkVec];
#pragma unroll
    for (int u = 0; u < kBlkU; ++u) {
      const long long e = base + ((long long)u * kBlkTpb + threadIdx.x) * kBlkVec;
#pragma unroll
      for (int j = 0; j < kBlkVec; ++j) {
        const bool ok = e + j >= e0 && e + j < e1;
        xv[u][j] = ok ? xs[(uint32_t)pw[u][j] >> kBlkRowBits] : 0.f;
      }
    }
#pragma unroll
    for (int u = 0; u < kBlkU; ++u) {
      const long long e = base + ((long long)u * kBlkTpb + threadIdx.x) * kBlkVec;
#pragma unroll
      for (int j = 0; j < kBlkVec; ++j) {
        if (e + j >= e0 && e + j < e1)
          __hip_atomic_fetch_add(&acc[(uint32_t)pw[u][j] & (kBlkRows - 1)], vv[u][j] * xv[u][j],
                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
  }
  __syncthreads();
  float* out = partial + (long long)s * part_stride + r0;
  for (int i = threadIdx.x; i < nrows; i += kBlkTpb) out[i] = acc[i];
}

// y[r] = ((partial_0[r] + partial_1[r]) + ...) + partial_{S-1}[r]
__global__ __launch_bounds__(256) void spmv_blocks_combine_kernel(
    const float* __restrict__ partial, long long part_stride, int S, int rows,
    float* __restrict__ y) {
  for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < rows;
       r += (long long)gridDim.x * blockDim.x) {
    float sum = partial[r];
    for (int s = 1; s < S; ++s) sum += partial[(long long)s * part_stride + r];
    y[r] = sum;
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {
inline size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

template <typename V>
hipError_t blk_scatter(const csr_args& A, const blk_geom& g, int* cursor, uint32_t* packed,
                       void* vals, unsigned grid, hipStream_t s) {
  blk_scatter_kernel<V><<<grid, 256, 0, s>>>(A.offsets, A.indices,
                                             static_cast<const V*>(A.values), g, cursor,
                                             packed, static_cast<V*>(vals));
  return hipGetLastError();
}

template <typename V>
hipError_t blk_execute(const blk_plan_state& st, const float* x, hipStream_t s) {
  spmv_blocks_kernel<V><<<(unsigned)((long long)st.nb * st.S), kBlkTpb, 0, s>>>(
      st.packed, static_cast<const V*>(st.values), st.stream_off, x, st.rows, st.S, st.W,
      st.partial, st.part_stride);
  return hipGetLastError();
}
}  // namespace

// Segments: a multiple of 8 (one XCD each), enough that W < 2^kBlkColBits.
int blk_segments(long long cols) {
  const long long per8 = 8LL << kBlkColBits;  // columns 8 segments can hold
  const long long groups = cols > 0 ? (cols + per8 - 1) / per8 : 1;
  return (int)(8 * groups);
}

hipError_t blk_plan_build(const csr_args& A, blk_plan_state* st, hipStream_t s) {
  blk_plan_state& P = *st;
  P.rows = A.rows;
  P.cols = A.cols;
  P.nnz = A.nnz;
  P.vtype = A.vtype;
  P.S = blk_segments(A.cols);
  P.W = (int)((A.cols + P.S - 1) / P.S);
  if (P.W < 1) P.W = 1;
  P.nb = (int)((A.rows + kBlkRows - 1) / kBlkRows);
  const int nch = (P.W + kBlkChunkCols - 1) / kBlkChunkCols;
  const long long nstreams = (long long)P.nb * P.S;
  const long long nbuckets = nstreams * nch;
  if (nbuckets + 1 >= (long long)INT32_MAX) return hipErrorInvalidValue;
  const int vbytes = A.vtype == LOOPS_VALUE_F32 ? 4 : 2;
  P.part_stride = (long long)((A.rows + 63) & ~63);

  // Plan memory: packed words and values (padded by one group of four),
  // stream offsets, partial outputs.
  const size_t n_pad = (size_t)A.nnz + kBlkVec;
  const size_t b_packed = al256(n_pad * sizeof(uint32_t));
  const size_t b_vals = al256(n_pad * vbytes);
  const size_t b_streams = al256((size_t)(nstreams + 1) * sizeof(int));
  const size_t b_part = al256((size_t)P.S * P.part_stride * sizeof(float));
  hipError_t e = hipMalloc(&P.mem, b_packed + b_vals + b_streams + b_part);
  if (e != hipSuccess) return e;
  char* m = static_cast<char*>(P.mem);
  P.packed = reinterpret_cast<uint32_t*>(m);
  P.values = m + b_packed;
  P.stream_off = reinterpret_cast<int*>(m + b_packed + b_vals);
  P.partial = reinterpret_cast<float*>(m + b_packed + b_vals + b_streams);
  P.device_bytes = b_packed + b_vals + b_streams + b_part;

  // Build scratch: bucket counts, bucket starts (+1), scan scratch.
  const size_t b_cnt = al256((size_t)nbuckets * sizeof(int));
  const size_t b_start = al256((size_t)(nbuckets + 1) * sizeof(int));
  const size_t b_scan = al256(exclusive_scan_scratch_bytes(nbuckets));
  char* sc = nullptr;
  if ((e = hipMalloc(&sc, b_cnt + b_start + b_scan)) != hipSuccess) return e;
  int* cnt = reinterpret_cast<int*>(sc);
  int* start = reinterpret_cast<int*>(sc + b_cnt);
  void* scan_scratch = sc + b_cnt + b_start;
  blk_geom g{A.rows, A.cols, A.nnz, P.S, P.W, nch, P.nb};
  long long grid = ((long long)A.nnz + 255) / 256;
  if (grid > 65536) grid = 65536;
  if (grid < 1) grid = 1;
  e = hipMemsetAsync(P.packed, 0, b_packed + b_vals, s);
  if (e == hipSuccess) e = hipMemsetAsync(cnt, 0, b_cnt, s);
  if (e == hipSuccess && A.nnz > 0) {
    blk_count_kernel<<<(unsigned)grid, 256, 0, s>>>(A.offsets, A.indices, g, cnt);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = launch_exclusive_scan(cnt, nbuckets, start, scan_scratch, s);
  if (e == hipSuccess) {
    blk_streams_kernel<<<(unsigned)((nstreams + 1 + 255) / 256), 256, 0, s>>>(
        start, (int)nstreams, nch, P.stream_off);
    e = hipGetLastError();
  }
  // the bucket starts become the scatter cursors
  if (e == hipSuccess && A.nnz > 0) {
    switch (A.vtype) {
      case LOOPS_VALUE_F32:
        e = blk_scatter<float>(A, g, start, P.packed, P.values, (unsigned)grid, s);
        break;
      case LOOPS_VALUE_BF16:
        e = blk_scatter<bf16_t>(A, g, start, P.packed, P.values, (unsigned)grid, s);
        break;
      case LOOPS_VALUE_F16:
        e = blk_scatter<f16_t>(A, g, start, P.packed, P.values, (unsigned)grid, s);
        break;
    }
  }
  const hipError_t e2 = hipStreamSynchronize(s);
  (void)hipFree(sc);
  return e != hipSuccess ? e : e2;
}

hipError_t blk_plan_execute(const blk_plan_state& st, const float* x, float* y,
                            hipEvent_t ev1, hipStream_t s) {
  hipError_t e = hipSuccess;
  if (st.rows == 0) return ev1 ? hipEventRecord(ev1, s) : hipSuccess;
  switch (st.vtype) {
    case LOOPS_VALUE_F32: e = blk_execute<float>(st, x, s); break;
    case LOOPS_VALUE_BF16: e = blk_execute<bf16_t>(st, x, s); break;
    case LOOPS_VALUE_F16: e = blk_execute<f16_t>(st, x, s); break;
  }
  if (e != hipSuccess) return e;
  if (ev1 && (e = hipEventRecord(ev1, s)) != hipSuccess) return e;
  long long grid = ((long long)st.rows + 255) / 256;
  if (grid > 8192) grid = 8192;
  spmv_blocks_combine_kernel<<<(unsigned)grid, 256, 0, s>>>(st.partial, st.part_stride, st.S,
                                                             st.rows, y);
  return hipGetLastError();
}

void blk_plan_release(blk_plan_state* st) {
  if (st->mem) (void)hipFree(st->mem);
  st->mem = nullptr;
}

}  // namespace loops_mi355x
