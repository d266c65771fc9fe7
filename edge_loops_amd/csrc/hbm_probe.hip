// HBM ceiling probes for the measurement harness (bench.py): how fast this
// GPU streams device memory with the access pattern the SpMV kernels use
// for their streamed arrays (16-byte non-temporal loads, grid-stride over
// whole kernels' worth of workgroups).  Not part of the C ABI of
// libloops_mi355x.so; built as libloops_hbm.so.
//
//   hbm_probe_read(src, bytes, grid, sink, stream)   reads `bytes` once
//   hbm_probe_copy(dst, src, bytes, grid, stream)    reads and writes them
//   hbm_probe_gather(ids, vals, n, x, grid, depth, sink, stream)
//       streams n column ids and values (16-byte non-temporal loads) and
//       gathers x at every id: a CSR matrix's nonzeros in CSR order without
//       the row walk or y, the least any CSR-order SpMV reads per nonzero
//   hbm_probe_gather_rows(ids, vals, n, B, ldb, c0, k, grid, depth, sink,
//                         stream)
//       for each of n nonzeros reads its column id and value and the k
//       floats B[id][c0 .. c0 + k) (16-byte loads, k / 4 lanes per row):
//       the least any CSR-order SpMM of those k columns reads per nonzero
//
// The read kernel folds every word into a per-lane xor and stores it only
// when it equals a constant no input produces in practice, so the loads
// cannot be dropped and nothing but one word per 2^32 lanes is written.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kTpb = 256;
constexpr int kDepth = 8;  // 16-byte loads in flight per lane

__global__ __launch_bounds__(kTpb) void hbm_read_kernel(const u32x4* __restrict__ src,
                                                        size_t n16, unsigned* __restrict__ sink) {
  const size_t stride = (size_t)gridDim.x * kTpb;
  size_t i = (size_t)blockIdx.x * kTpb + threadIdx.x;
  unsigned acc = 0;
  for (; i + (kDepth - 1) * stride < n16; i += kDepth * stride) {
    u32x4 v[kDepth];
#pragma unroll
    for (int j = 0; j < kDepth; ++j) v[j] = __builtin_nontemporal_load(src + i + j * stride);
#pragma unroll
    for (int j = 0; j < kDepth; ++j) acc ^= v[j].x ^ v[j].y ^ v[j].z ^ v[j].w;
  }
  for (; i < n16; i += stride) {
    const u32x4 v = __builtin_nontemporal_load(src + i);
    acc ^= v.x ^ v.y ^ v.z ^ v.w;
  }
  if (acc == 0x9e3779b9u) sink[threadIdx.x] = acc;
}

__global__ __launch_bounds__(kTpb) void hbm_copy_kernel(u32x4* __restrict__ dst,
                                                        const u32x4* __restrict__ src,
                                                        size_t n16) {
  const size_t stride = (size_t)gridDim.x * kTpb;
  size_t i = (size_t)blockIdx.x * kTpb + threadIdx.x;
  for (; i + (kDepth - 1) * stride < n16; i += kDepth * stride) {
    u32x4 v[kDepth];
#pragma unroll
    for (int j = 0; j < kDepth; ++j) v[j] = __builtin_nontemporal_load(src + i + j * stride);
#pragma unroll
    for (int j = 0; j < kDepth; ++j) __builtin_nontemporal_store(v[j], dst + i + j * stride);
  }
  for (; i < n16; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int kD>  // 16-byte loads of ids and of values in flight per lane: 2 kD
__global__ __launch_bounds__(kTpb) void hbm_gather_kernel(const i32x4* __restrict__ ids,
                                                          const f32x4* __restrict__ vals,
                                                          size_t n4, const int* __restrict__ tid,
                                                          const float* __restrict__ tval, int ntail,
                                                          const float* __restrict__ x,
                                                          unsigned* __restrict__ sink) {
  const size_t stride = (size_t)gridDim.x * kTpb;
  size_t i = (size_t)blockIdx.x * kTpb + threadIdx.x;
  float acc = 0.f;
  for (; i + (kD - 1) * stride < n4; i += kD * stride) {
    i32x4 c[kD];
    f32x4 v[kD];
#pragma unroll
    for (int j = 0; j < kD; ++j) {
      c[j] = __builtin_nontemporal_load(ids + i + j * stride);
      v[j] = __builtin_nontemporal_load(vals + i + j * stride);
    }
#pragma unroll
    for (int j = 0; j < kD; ++j)
      acc += v[j].x * x[c[j].x] + v[j].y * x[c[j].y] + v[j].z * x[c[j].z] + v[j].w * x[c[j].w];
  }
  for (; i < n4; i += stride) {
    const i32x4 c = __builtin_nontemporal_load(ids + i);
    const f32x4 v = __builtin_nontemporal_load(vals + i);
    acc += v.x * x[c.x] + v.y * x[c.y] + v.z * x[c.z] + v.w * x[c.w];
  }
  if (blockIdx.x == 0 && (int)threadIdx.x < ntail) acc += tval[threadIdx.x] * x[tid[threadIdx.x]];
  if (__float_as_uint(acc) == 0x9e3779b9u) sink[threadIdx.x] = 1u;
}

// G lanes per nonzero (4 columns each), D nonzeros in flight per lane group.
template <int G, int D>
__global__ __launch_bounds__(kTpb) void hbm_gather_rows_kernel(
    const int* __restrict__ ids, const float* __restrict__ vals, size_t n,
    const float* __restrict__ B, long long ldb, int c0, unsigned* __restrict__ sink) {
  const size_t groups = (size_t)gridDim.x * (kTpb / G);
  const int sub = threadIdx.x % G;
  size_t nz = (size_t)blockIdx.x * (kTpb / G) + threadIdx.x / G;
  const float* __restrict__ Bc = B + c0 + 4 * sub;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (; nz + (D - 1) * groups < n; nz += D * groups) {
    int c[D];
    float v[D];
    f32x4 b[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      c[j] = ids[nz + j * groups];
      v[j] = vals[nz + j * groups];
    }
#pragma unroll
    for (int j = 0; j < D; ++j) b[j] = *reinterpret_cast<const f32x4*>(Bc + (long long)c[j] * ldb);
#pragma unroll
    for (int j = 0; j < D; ++j) acc += v[j] * b[j];
  }
  for (; nz < n; nz += groups)
    acc += vals[nz] * *reinterpret_cast<const f32x4*>(Bc + (long long)ids[nz] * ldb);
  if (__float_as_uint(acc.x + acc.y + acc.z + acc.w) == 0x9e3779b9u) sink[threadIdx.x] = 1u;
}

}  // namespace

extern "C" {

// bytes must be a multiple of 16 and the pointers 16-byte aligned; grid > 0;
// sink: device memory of at least 256 unsigned words.
int hbm_probe_read(const void* src, size_t bytes, int grid, unsigned* sink, hipStream_t stream) {
  if (!src || !sink || grid <= 0 || bytes % 16 != 0 ||
      reinterpret_cast<uintptr_t>(src) % 16 != 0)
    return (int)hipErrorInvalidValue;
  hbm_read_kernel<<<grid, kTpb, 0, stream>>>(static_cast<const u32x4*>(src), bytes / 16, sink);
  return (int)hipGetLastError();
}

int hbm_probe_copy(void* dst, const void* src, size_t bytes, int grid, hipStream_t stream) {
  if (!src || !dst || grid <= 0 || bytes % 16 != 0 ||
      reinterpret_cast<uintptr_t>(src) % 16 != 0 || reinterpret_cast<uintptr_t>(dst) % 16 != 0)
    return (int)hipErrorInvalidValue;
  hbm_copy_kernel<<<grid, kTpb, 0, stream>>>(static_cast<u32x4*>(dst),
                                             static_cast<const u32x4*>(src), bytes / 16);
  return (int)hipGetLastError();
}

// ids, vals: 16-byte aligned, n entries; every id indexes x (the caller's
// CSR column ids, which it has checked); grid > 0; depth 1, 2, 4 or 8
// (pairs of 16-byte loads in flight per lane, 4 gathers each); sink as for
// the read.
int hbm_probe_gather(const int* ids, const float* vals, size_t n, const float* x, int grid,
                     int depth, unsigned* sink, hipStream_t stream) {
  if (!ids || !vals || !x || !sink || grid <= 0 ||
      reinterpret_cast<uintptr_t>(ids) % 16 != 0 || reinterpret_cast<uintptr_t>(vals) % 16 != 0)
    return (int)hipErrorInvalidValue;
  const size_t n4 = n / 4;
  const i32x4* i4 = reinterpret_cast<const i32x4*>(ids);
  const f32x4* v4 = reinterpret_cast<const f32x4*>(vals);
  const int* ti = ids + 4 * n4;
  const float* tv = vals + 4 * n4;
  const int nt = (int)(n - 4 * n4);
  switch (depth) {
    case 1: hbm_gather_kernel<1><<<grid, kTpb, 0, stream>>>(i4, v4, n4, ti, tv, nt, x, sink); break;
    case 2: hbm_gather_kernel<2><<<grid, kTpb, 0, stream>>>(i4, v4, n4, ti, tv, nt, x, sink); break;
    case 4: hbm_gather_kernel<4><<<grid, kTpb, 0, stream>>>(i4, v4, n4, ti, tv, nt, x, sink); break;
    case 8: hbm_gather_kernel<8><<<grid, kTpb, 0, stream>>>(i4, v4, n4, ti, tv, nt, x, sink); break;
    default: return (int)hipErrorInvalidValue;
  }
  return (int)hipGetLastError();
}

// ids, vals: n entries, every id a row of B; B, ldb and c0 such that every
// B + id * ldb + c0 is 16-byte aligned; k 4, 8, 16, 32 or 64 columns;
// depth 2, 4 or 8 nonzeros in flight per lane group; grid > 0.
int hbm_probe_gather_rows(const int* ids, const float* vals, size_t n, const float* B,
                          long long ldb, int c0, int k, int grid, int depth, unsigned* sink,
                          hipStream_t stream) {
  if (!ids || !vals || !B || !sink || grid <= 0 || c0 < 0 || ldb % 4 != 0 || c0 % 4 != 0 ||
      c0 + k > ldb || reinterpret_cast<uintptr_t>(B) % 16 != 0)
    return (int)hipErrorInvalidValue;
#define LOOPS_ROWS_PROBE(G, D)                                                          \
  if (k == 4 * G && depth == D) {                                                       \
    hbm_gather_rows_kernel<G, D><<<grid, kTpb, 0, stream>>>(ids, vals, n, B, ldb, c0, sink); \
    return (int)hipGetLastError();                                                      \
  }
#define LOOPS_ROWS_PROBE_G(G) LOOPS_ROWS_PROBE(G, 2) LOOPS_ROWS_PROBE(G, 4) LOOPS_ROWS_PROBE(G, 8)
  LOOPS_ROWS_PROBE_G(1) LOOPS_ROWS_PROBE_G(2) LOOPS_ROWS_PROBE_G(4) LOOPS_ROWS_PROBE_G(8)
  LOOPS_ROWS_PROBE_G(16)
#undef LOOPS_ROWS_PROBE_G
#undef LOOPS_ROWS_PROBE
  return (int)hipErrorInvalidValue;
}

}  // extern "C"
