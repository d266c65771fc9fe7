// Device helpers shared by the gfx950 kernels of libloops_mi355x.so:
// value-storage conversions, streaming vector loads and wavefront /
// workgroup segmented scans.  Everything here is written for 64-lane
// wavefronts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace loops_mi355x {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// bfloat16 storage: the upper 16 bits of an IEEE binary32.
struct bf16_t {
  uint16_t bits;
};
using f16_t = _Float16;

template <typename V>
struct value_traits;

template <>
struct value_traits<float> {
  static constexpr int bytes = 4;
  __device__ static __forceinline__ float load1(const float* p) {
    return __builtin_nontemporal_load(p);
  }
  // Four consecutive values from a 16-byte aligned address.
  __device__ static __forceinline__ f32x4 load4(const float* p) {
    return __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(p));
  }
};

template <>
struct value_traits<bf16_t> {
  static constexpr int bytes = 2;
  __device__ static __forceinline__ float widen(uint32_t b) {
    return __uint_as_float(b << 16);
  }
  __device__ static __forceinline__ float load1(const bf16_t* p) {
    uint16_t b = __builtin_nontemporal_load(reinterpret_cast<const uint16_t*>(p));
    return widen(b);
  }
  // Four consecutive values from an 8-byte aligned address.
  __device__ static __forceinline__ f32x4 load4(const bf16_t* p) {
    u32x2 w = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(p));
    f32x4 r;
    r.x = __uint_as_float(w.x << 16);
    r.y = __uint_as_float(w.x & 0xffff0000u);
    r.z = __uint_as_float(w.y << 16);
    r.w = __uint_as_float(w.y & 0xffff0000u);
    return r;
  }
};

template <>
struct value_traits<f16_t> {
  static constexpr int bytes = 2;
  __device__ static __forceinline__ float load1(const f16_t* p) {
    uint16_t b = __builtin_nontemporal_load(reinterpret_cast<const uint16_t*>(p));
    return static_cast<float>(__builtin_bit_cast(f16_t, b));
  }
  __device__ static __forceinline__ f32x4 load4(const f16_t* p) {
    u32x2 w = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(p));
    f32x4 r;
    r.x = static_cast<float>(__builtin_bit_cast(f16_t, (uint16_t)(w.x & 0xffffu)));
    r.y = static_cast<float>(__builtin_bit_cast(f16_t, (uint16_t)(w.x >> 16)));
    r.z = static_cast<float>(__builtin_bit_cast(f16_t, (uint16_t)(w.y & 0xffffu)));
    r.w = static_cast<float>(__builtin_bit_cast(f16_t, (uint16_t)(w.y >> 16)));
    return r;
  }
};

__device__ __forceinline__ int ld_idx(const int* p) {
  return __builtin_nontemporal_load(p);
}
__device__ __forceinline__ i32x4 ld_idx4(const int* p) {
  return __builtin_nontemporal_load(reinterpret_cast<const i32x4*>(p));
}

// Segmented inclusive scan over groups of W lanes (W <= 64, power of two).
// Element = (flag, value); combine(a, b) = (a.f | b.f, b.f ? b.v : a.v + b.v).
// On return v holds the inclusive value and f the inclusive flag.
template <int W>
__device__ __forceinline__ void group_seg_scan(float& v, int& f) {
  const int l = threadIdx.x & (W - 1);
#pragma unroll
  for (int d = 1; d < W; d <<= 1) {
    float vu = __shfl_up(v, d, W);
    int fu = __shfl_up(f, d, W);
    if (l >= d) {
      if (!f) v = vu + v;
      f |= fu;
    }
  }
}

// Segmented inclusive scan over a whole workgroup of TPB lanes (TPB a
// multiple of 64).  s_v / s_f hold TPB/64 entries.  Contains two barriers;
// every lane of the block must call it.  Also returns, through pv/pf, the
// combined value of all earlier waves (the identity (0, 0) for wave 0).
template <int TPB>
__device__ __forceinline__ void block_seg_scan(float& v, int& f, float* s_v,
                                               int* s_f, float& pv, int& pf) {
  constexpr int NW = TPB / 64;
  group_seg_scan<64>(v, f);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 63) {
    s_v[wave] = v;
    s_f[wave] = f;
  }
  __syncthreads();
  pv = 0.f;
  pf = 0;
  for (int w = 0; w < wave && w < NW; ++w) {
    float wv = s_v[w];
    int wf = s_f[w];
    pv = wf ? wv : pv + wv;
    pf |= wf;
  }
  if (!f) v = pv + v;
  f |= pf;
  __syncthreads();
}

}  // namespace loops_mi355x
