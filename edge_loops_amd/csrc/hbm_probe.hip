// HBM ceiling probes for the measurement harness (bench.py): how fast this
// GPU streams device memory with the access pattern the SpMV kernels use
// for their streamed arrays (16-byte non-temporal loads, grid-stride over
// whole kernels' worth of workgroups).  Not part of the C ABI of
// libloops_mi355x.so; built as libloops_hbm.so.
//
//   hbm_probe_read(src, bytes, grid, sink, stream)   reads `bytes` once
//   hbm_probe_copy(dst, src, bytes, grid, stream)    reads and writes them
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

}  // extern "C"
