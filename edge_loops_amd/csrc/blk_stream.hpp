// Row-block SpMV streams: the device code of one (block, segment) stream of
// a row-block plan and of the combine of its partials, shared by the
// plan's own kernels (spmv_blocks.hip) and by the batch kernels
// (batch_kernels.hip), so that a plan gives the same bits in both.
//
// A stream holds the rounds of one (block, segment) pair: at most
// kBlkRound entries each, in column order, stored lane-major per
// 512-entry slab.  A workgroup of kBlkTpb lanes keeps the y of the block's
// rows in LDS and, per round, multiplies in column order into LDS slots
// (phase 1), then sums runs of equal rows over each lane's 8 slots with a
// segmented workgroup scan (phase 2) and adds each run's total to its row,
// one writer per row per round.  At the end it writes the rows' partial
// sums for its segment.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "device_common.hpp"
#include "loops_internal.hpp"

namespace loops_mi355x {

constexpr int kBlkTpb = 512;
constexpr int kBlkIpt = kBlkRound / kBlkTpb;  // entries per lane per phase (8)
constexpr int kBlkSlab = 64 * kBlkIpt;        // entries of one wavefront per round
static_assert(kBlkIpt == 8 && kBlkIpt * kBlkTpb == kBlkRound,
              "a round is one pass of the workgroup, 8 slots (16 bytes of rows) per lane");

// What executing a plan reads and writes besides x and y (the device
// arrays of blk_plan_state).
struct blk_exec_args {
  const uint32_t* packed;
  const void* values;
  const uint16_t* srow;
  const int* stream_round;
  const int* round_off;
  const int* round_base;
  const int* blk_start;
  const int* blk_order;
  const int* rstart;
  const int* rlist;
  const int* vrow;
  float* partial;
  long long part_stride;
  int nb, S, W, rows;
};

// y: where a direct plan's streams write (the partials of the others).
inline blk_exec_args blk_args_of(const blk_plan_state& st, float* y) {
  blk_exec_args a;
  a.packed = st.packed;
  a.values = st.values;
  a.srow = st.srow;
  a.stream_round = st.stream_round;
  a.round_off = st.round_off;
  a.round_base = st.round_base;
  a.blk_start = st.blk_start;
  a.blk_order = st.blk_order;
  a.rstart = st.rstart;
  a.rlist = st.rlist;
  a.vrow = st.vrow;
  a.partial = st.direct ? y : st.partial;
  a.part_stride = st.part_stride;
  a.nb = st.nb;
  a.S = st.S;
  a.W = st.W;
  a.rows = st.rows;
  return a;
}

// A load of a value the whole wavefront shares, from plan arrays that no
// kernel writes while streams run: through the constant address space, so
// that it is a scalar load (served by the scalar cache, off the vector memory
// path the x gathers keep busy; a vector load here sits in the round's
// critical path and cost the C2 stream kernel 15 %).
template <typename T>
__device__ __forceinline__ T blk_ld_uniform(const T* p) {
  return *(const __attribute__((address_space(4))) T*)p;
}

// LDS of one stream's workgroup: the block's y and a round's products.
struct blk_smem {
  float acc[kBlkRows];
  __align__(16) float stage[kBlkRound];
  __align__(16) float s_v[kBlkTpb / 64];
  __align__(16) int s_f[kBlkTpb / 64];
};

typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

// What a lane reads of one round: its phase-1 entries (wavefront w: entries
// 512 w .. 512 w + 511, lane l: 64 j + l, stored lane-major so that the
// lane's 8 are consecutive) and the rows of its phase-2 slots 8 t .. 8 t + 7;
// the wave also reads the rows of the slots just before and after its 512.
template <typename V>
struct blk_round_regs {
  int n;
  uint32_t pw[kBlkIpt];
  uint32_t vw[kBlkIpt];  // raw value words (blk_word_value)
  u16x8 rows;
  int edge_prev, edge_next;  // wave-uniform: the rows of the slots just before / after the wave's
};

// Values travel as raw 32-bit words and are converted where they are used,
// so no load result is waited for inside the guarded load block.  A lane's
// 8 16-bit values are 4 words; value j is half j & 1 of word j / 2.
template <typename V>
__device__ __forceinline__ float blk_word_value(uint32_t w, int j) {
  if constexpr (std::is_same<V, float>::value) {
    return __uint_as_float(w);
  } else {
    const uint32_t h = (j & 1) ? (w >> 16) : (w & 0xffffu);
    if constexpr (std::is_same<V, bf16_t>::value) return __uint_as_float(h << 16);
    else return static_cast<float>(__builtin_bit_cast(f16_t, (uint16_t)h));
  }
}

template <typename V>
__device__ __forceinline__ void blk_load_round(int r, const uint32_t* __restrict__ packed,
                                               const V* __restrict__ vals,
                                               const uint16_t* __restrict__ srow,
                                               const int* __restrict__ round_off,
                                               const int* __restrict__ round_base, int lane,
                                               int wave, int k0, blk_round_regs<V>& g) {
  g.n = blk_ld_uniform(round_off + r + 1) - blk_ld_uniform(round_off + r);
#pragma unroll
  for (int j = 0; j < kBlkIpt; ++j) {
    g.pw[j] = 0;
    g.vw[j] = 0;
  }
  // The wave's slab exists when its first entry does; positions past the
  // round's end inside it are zero padding.
  if (wave * kBlkSlab < g.n) {
    const long long s0 = ((long long)blk_ld_uniform(round_base + r) + wave) * kBlkSlab;
    const long long p = s0 + lane * 4, p2 = p + kBlkSlab / 2;
    const i32x4 c0 = __builtin_nontemporal_load(reinterpret_cast<const i32x4*>(packed + p));
    const i32x4 c1 = __builtin_nontemporal_load(reinterpret_cast<const i32x4*>(packed + p2));
    g.pw[0] = c0.x; g.pw[1] = c0.y; g.pw[2] = c0.z; g.pw[3] = c0.w;
    g.pw[4] = c1.x; g.pw[5] = c1.y; g.pw[6] = c1.z; g.pw[7] = c1.w;
    if constexpr (std::is_same<V, float>::value) {
      const i32x4 v0 = __builtin_nontemporal_load(reinterpret_cast<const i32x4*>(vals + p));
      const i32x4 v1 = __builtin_nontemporal_load(reinterpret_cast<const i32x4*>(vals + p2));
      g.vw[0] = v0.x; g.vw[1] = v0.y; g.vw[2] = v0.z; g.vw[3] = v0.w;
      g.vw[4] = v1.x; g.vw[5] = v1.y; g.vw[6] = v1.z; g.vw[7] = v1.w;
    } else {
      const i32x4 v = __builtin_nontemporal_load(
          reinterpret_cast<const i32x4*>(vals + s0 + (long long)lane * kBlkIpt));
      g.vw[0] = g.vw[1] = v.x; g.vw[2] = g.vw[3] = v.y;
      g.vw[4] = g.vw[5] = v.z; g.vw[6] = g.vw[7] = v.w;
    }
  }
  const uint16_t* rs = srow + (long long)r * kBlkRound;
  g.rows = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
  if (k0 < g.n) g.rows = __builtin_nontemporal_load(reinterpret_cast<const u16x8*>(rs + k0));
  // The neighbours' rows inside the wave come from lane shuffles (phase 2);
  // only the two slots at the wave's edges are loaded, as aligned words at
  // wave-uniform addresses.
  const uint32_t* rw = reinterpret_cast<const uint32_t*>(rs);
  const int base = __builtin_amdgcn_readfirstlane(wave) * kBlkSlab;
  g.edge_prev = -1;
  g.edge_next = -1;
  if (base > 0 && base < g.n) g.edge_prev = (int)(blk_ld_uniform(rw + ((base - 2) >> 1)) >> 16);
  if (base + kBlkSlab < g.n)
    g.edge_next = (int)(blk_ld_uniform(rw + ((base + kBlkSlab) >> 1)) & 0xffffu);
}

// Loads round r of a stream into g; live == false (a round past the
// stream's end) gives an empty round (n = 0).  Every vector load is issued
// whatever the round holds, so each round issues the same vector loads and
// the compiler's wait counts stay exact across the loop (a conditional load,
// or a select on a loaded value, makes it wait for every load in flight,
// the prefetched ones included).  Wavefronts and lanes with nothing to load
// read the plan's first slab and round 0's slot rows instead: valid
// addresses, shared by all of them (so they hit in cache), holding column
// words of the plan (so the x reads they lead to stay inside the segment);
// what they read is never stored or summed (slots past n are not used).
template <typename V>
__device__ __forceinline__ void blk_load_round_all(int r, bool live,
                                               const uint32_t* __restrict__ packed,
                                               const V* __restrict__ vals,
                                               const uint16_t* __restrict__ srow,
                                               const int* __restrict__ round_off,
                                               const int* __restrict__ round_base, int lane,
                                               int wave, int k0, blk_round_regs<V>& g) {
  g.n = live ? blk_ld_uniform(round_off + r + 1) - blk_ld_uniform(round_off + r) : 0;
  // The wave's slab exists when its first entry does; positions past the
  // round's end inside it are zero padding.
  const bool has = wave * kBlkSlab < g.n;
  const long long s0 = has ? ((long long)blk_ld_uniform(round_base + r) + wave) * kBlkSlab : 0;
  const long long p = s0 + lane * 4, p2 = p + kBlkSlab / 2;
  const i32x4 c0 = __builtin_nontemporal_load(reinterpret_cast<const i32x4*>(packed + p));
  const i32x4 c1 = __builtin_nontemporal_load(reinterpret_cast<const i32x4*>(packed + p2));
  g.pw[0] = c0.x; g.pw[1] = c0.y; g.pw[2] = c0.z; g.pw[3] = c0.w;
  g.pw[4] = c1.x; g.pw[5] = c1.y; g.pw[6] = c1.z; g.pw[7] = c1.w;
  if constexpr (std::is_same<V, float>::value) {
    const i32x4 v0 = __builtin_nontemporal_load(reinterpret_cast<const i32x4*>(vals + p));
    const i32x4 v1 = __builtin_nontemporal_load(reinterpret_cast<const i32x4*>(vals + p2));
    g.vw[0] = v0.x; g.vw[1] = v0.y; g.vw[2] = v0.z; g.vw[3] = v0.w;
    g.vw[4] = v1.x; g.vw[5] = v1.y; g.vw[6] = v1.z; g.vw[7] = v1.w;
  } else {
    const i32x4 v = __builtin_nontemporal_load(
        reinterpret_cast<const i32x4*>(vals + s0 + (long long)lane * kBlkIpt));
    g.vw[0] = g.vw[1] = v.x; g.vw[2] = g.vw[3] = v.y;
    g.vw[4] = g.vw[5] = v.z; g.vw[6] = g.vw[7] = v.w;
  }
  const uint16_t* rs = srow + (k0 < g.n ? (long long)r * kBlkRound + k0 : 0);
  g.rows = __builtin_nontemporal_load(reinterpret_cast<const u16x8*>(rs));
  // The neighbours' rows inside the wave come from lane shuffles (phase 2);
  // only the two slots at the wave's edges are loaded, as aligned words at
  // wave-uniform addresses.
  const uint32_t* rw = reinterpret_cast<const uint32_t*>(srow + (long long)r * kBlkRound);
  const int base = __builtin_amdgcn_readfirstlane(wave) * kBlkSlab;
  g.edge_prev = -1;
  g.edge_next = -1;
  if (base > 0 && base < g.n) g.edge_prev = (int)(blk_ld_uniform(rw + ((base - 2) >> 1)) >> 16);
  if (base + kBlkSlab < g.n)
    g.edge_next = (int)(blk_ld_uniform(rw + ((base + kBlkSlab) >> 1)) & 0xffffu);
}

// Segmented inclusive scan over a wavefront with DPP lane moves: shifts by
// 1, 2, 4 and 8 inside rows of 16 lanes, then each row takes the total of
// the rows before it (row_bcast 15 into rows 1 and 3, row_bcast 31 into
// rows 2 and 3).  A lane whose source is outside the move receives (0, 0),
// the identity of the combination (v, f) + (v', f') = (f' ? v' : v + v',
// f | f').
__device__ __forceinline__ void blk_wave_scan(float& v, int& f) {
#define BLK_SCAN_STEP(CTRL, ROWS)                                                          \
  {                                                                                        \
    const float vu =                                                                       \
        __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xf, \
                                                   false));                                \
    const int fu = __builtin_amdgcn_update_dpp(0, f, CTRL, ROWS, 0xf, false);              \
    if (!f) v = vu + v;                                                                    \
    f |= fu;                                                                               \
  }
  BLK_SCAN_STEP(0x111, 0xf)  // row_shr:1
  BLK_SCAN_STEP(0x112, 0xf)  // row_shr:2
  BLK_SCAN_STEP(0x114, 0xf)  // row_shr:4
  BLK_SCAN_STEP(0x118, 0xf)  // row_shr:8
  BLK_SCAN_STEP(0x142, 0xa)  // row_bcast:15 into rows 1 and 3
  BLK_SCAN_STEP(0x143, 0xc)  // row_bcast:31 into rows 2 and 3
#undef BLK_SCAN_STEP
}

// Segmented inclusive scan over the workgroup (block_seg_scan's result):
// wavefront scans, then each lane combines the totals of the wavefronts
// before its own, read with 16-byte LDS reads.  One barrier; every lane
// must call it.  The totals are next written in the following round's scan,
// after that round's phase-1 barrier, which every wave reaches only once it
// has read them here.
__device__ __forceinline__ void blk_block_scan(float& v, int& f, float* s_v, int* s_f,
                                               float& pv, int& pf) {
  constexpr int NW = kBlkTpb / 64;
  static_assert(NW == 4 || NW == 8, "four or eight wavefront totals, 16-byte reads");
  blk_wave_scan(v, f);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 63) {
    s_v[wave] = v;
    s_f[wave] = f;
  }
  __syncthreads();
  float wv[NW];
  int wf[NW];
#pragma unroll
  for (int h = 0; h < NW; h += 4) {
    const f32x4 vv = *reinterpret_cast<const f32x4*>(s_v + h);
    const i32x4 ff = *reinterpret_cast<const i32x4*>(s_f + h);
    wv[h] = vv.x; wv[h + 1] = vv.y; wv[h + 2] = vv.z; wv[h + 3] = vv.w;
    wf[h] = ff.x; wf[h + 1] = ff.y; wf[h + 2] = ff.z; wf[h + 3] = ff.w;
  }
  pv = 0.f;
  pf = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    if (w < wave) {
      pv = wf[w] ? wv[w] : pv + wv[w];
      pf |= wf[w];
    }
  }
  if (!f) v = pv + v;
  f |= pf;
}

// Stream `piece` of a plan (0 <= piece < nb * S): segment piece % S of
// block blk_order[piece / S], so consecutive pieces take consecutive
// segments.  Every lane of the workgroup calls it; the caller separates
// two calls on one workgroup with a barrier (the LDS is reused).
template <typename V, bool kAhead>
__device__ __forceinline__ void blk_stream(const blk_exec_args& a, const float* __restrict__ x,
                                           int piece, blk_smem& sm) {
  const uint32_t* __restrict__ packed = a.packed;
  const V* __restrict__ vals = static_cast<const V*>(a.values);
  const uint16_t* __restrict__ srow = a.srow;
  const int* __restrict__ round_off = a.round_off;
  const int* __restrict__ round_base = a.round_base;
  float* acc = sm.acc;
  float* stage = sm.stage;
  const int S = a.S;
  const int s = piece % S;
  const int b = blk_ld_uniform(a.blk_order + piece / S);
  const int q = b * S + s;  // stream
  // stream rows: the compressed list positions rs .. rs + nrows - 1, or the
  // block's rows r0 .. r0 + nrows - 1 (<= kBlkRows either way)
  int r0 = 0, rs = 0, nrows;
  if (a.rlist) {
    const int* rp = a.rstart + (long long)s * (a.nb + 1) + b;
    rs = blk_ld_uniform(rp);
    nrows = blk_ld_uniform(rp + 1) - rs;
  } else {
    r0 = blk_ld_uniform(a.blk_start + b);
    nrows = blk_ld_uniform(a.blk_start + b + 1) - r0;
  }
  const float* __restrict__ xs = x + (long long)s * a.W;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int k0 = threadIdx.x * kBlkIpt;
  const int rb = blk_ld_uniform(a.stream_round + q), re = blk_ld_uniform(a.stream_round + q + 1);
  // Phase 2 of a round (after the barrier that follows its stage writes):
  // runs of equal rows over slots k0 .. k0 + m - 1.  prev_row / next_row:
  // the rows of slots k0 - 1 and k0 + kBlkIpt (-1 outside the round), from
  // the neighbouring lanes or the wave's edge slots.
  auto phase2 = [&](const blk_round_regs<V>& cur) {
    const int n = cur.n;
    const int m = n - k0 < 0 ? 0 : (n - k0 > kBlkIpt ? kBlkIpt : n - k0);
    const int up_row = __shfl_up((int)cur.rows[kBlkIpt - 1], 1, 64);
    const int down_row = __shfl_down((int)cur.rows[0], 1, 64);
    const int prev_row = k0 == 0 ? -1 : (lane == 0 ? cur.edge_prev : up_row);
    const int next_row = k0 + kBlkIpt >= n ? -1 : (lane == 63 ? cur.edge_next : down_row);
    float v = 0.f, first_sum = 0.f;
    int f = 1, first_row = -1, row = cur.rows[0];
    bool internal = false;
    // the lane's 8 slots with two 16-byte LDS reads (slots past the round's
    // end hold stale products and are not used)
    const f32x4 sa = *reinterpret_cast<const f32x4*>(stage + k0);
    const f32x4 sb = *reinterpret_cast<const f32x4*>(stage + k0 + 4);
    const float sv[kBlkIpt] = {sa.x, sa.y, sa.z, sa.w, sb.x, sb.y, sb.z, sb.w};
    if (m > 0) {
      v = sv[0];
#pragma unroll
      for (int i = 1; i < kBlkIpt; ++i) {
        if (i < m) {
          const float p = sv[i];
          const int ri = cur.rows[i];
          if (ri != row) {
            if (!internal) {
              first_sum = v;
              first_row = row;
              internal = true;
            } else {
              acc[row] += v;  // a run that starts and ends in this lane
            }
            row = ri;
            v = p;
          } else {
            v += p;
          }
        }
      }
      f = internal || prev_row != (int)cur.rows[0];
    }
    const float v_own = v;
    float pv;
    int pf;
    blk_block_scan(v, f, sm.s_v, sm.s_f, pv, pf);
    // exclusive: the running sum of the previous lane's trailing run
    const float up = __shfl_up(v, 1, 64);
    const float ex = lane == 0 ? pv : up;
    if (m > 0) {
      const bool cont = k0 > 0 && prev_row == (int)cur.rows[0];
      const float cin = cont ? ex : 0.f;
      if (internal) acc[first_row] += cin + first_sum;
      if (k0 + m >= n || next_row != row) acc[row] += (internal ? 0.f : cin) + v_own;
    }
  };
  if constexpr (!kAhead) {
    // One round in flight: a round's x reads and the next round's loads go
    // out together and are waited for before its stage writes.
    blk_round_regs<V> cur;
    if (rb < re)
      blk_load_round<V>(rb, packed, vals, srow, round_off, round_base, lane, wave, k0, cur);
    // the accumulator is cleared while the first round's loads are in flight
    for (int i = threadIdx.x; i < nrows; i += kBlkTpb) acc[i] = 0.f;
    for (int r = rb; r < re; ++r) {
      const int n = cur.n;
      // Phase 1: products in column order to their slots; the next round's
      // loads go out behind this round's x reads.
      float xv[kBlkIpt];
#pragma unroll
      for (int j = 0; j < kBlkIpt; ++j) {
        const int i = wave * 64 * kBlkIpt + j * 64 + lane;
        xv[j] = i < n ? xs[cur.pw[j] >> kBlkSlotBits] : 0.f;
      }
      blk_round_regs<V> nxt;
      if (r + 1 < re)
        blk_load_round<V>(r + 1, packed, vals, srow, round_off, round_base, lane, wave, k0, nxt);
#pragma unroll
      for (int j = 0; j < kBlkIpt; ++j) {
        const int i = wave * 64 * kBlkIpt + j * 64 + lane;
        if (i < n)
          stage[cur.pw[j] & (kBlkRound - 1)] = blk_word_value<V>(cur.vw[j], j) * xv[j];
      }
      __syncthreads();
      phase2(cur);
      if (r + 1 < re) cur = nxt;
    }
  } else {
    // Two rounds in flight: right after a round's stage writes, the next
    // round's x reads and the loads of the round after it go out, so they
    // are in flight through this round's phase 2 (every vector load issued
    // unconditionally, blk_load_round_all, so the waits stay exact).
    // (entries past a round's end have column words of the plan and read
    // inside the segment; their products are not stored)
    auto gather = [&](const blk_round_regs<V>& g, float* xv) {
#pragma unroll
      for (int j = 0; j < kBlkIpt; ++j) xv[j] = xs[g.pw[j] >> kBlkSlotBits];
    };
    blk_round_regs<V> cur, nxt;
    float xv[kBlkIpt];
    if (rb < re)
      blk_load_round_all<V>(rb, true, packed, vals, srow, round_off, round_base, lane, wave, k0,
                            cur);
    for (int i = threadIdx.x; i < nrows; i += kBlkTpb) acc[i] = 0.f;
    if (rb < re) {
      gather(cur, xv);
      blk_load_round_all<V>(rb + 1, rb + 1 < re, packed, vals, srow, round_off, round_base, lane,
                            wave, k0, nxt);
    }
    for (int r = rb; r < re; ++r) {
      const int n = cur.n;
#pragma unroll
      for (int j = 0; j < kBlkIpt; ++j) {
        const int i = wave * 64 * kBlkIpt + j * 64 + lane;
        if (i < n)
          stage[cur.pw[j] & (kBlkRound - 1)] = blk_word_value<V>(cur.vw[j], j) * xv[j];
      }
      blk_round_regs<V> nn;
      gather(nxt, xv);
      blk_load_round_all<V>(r + 2, r + 2 < re, packed, vals, srow, round_off, round_base, lane,
                            wave, k0, nn);
      __syncthreads();
      phase2(cur);
      cur = nxt;
      nxt = nn;
    }
  }
  __syncthreads();
  float* out = a.partial + (long long)s * a.part_stride;
  if (a.rlist) {
    const int* __restrict__ rl = a.rlist + rs;
    for (int i = threadIdx.x; i < nrows; i += kBlkTpb)
      __builtin_nontemporal_store(acc[i], out + rl[i]);
  } else {
    for (int i = threadIdx.x; i < nrows; i += kBlkTpb)
      __builtin_nontemporal_store(acc[i], out + r0 + i);
  }
}

// Which stream loop runs a plan: two rounds in flight for plans of one
// column segment (the plans cut for batches), one round otherwise.  The
// 2M-row matrix's plan (4 segments, x slices held in L2, 20 rounds per
// stream) ran 129 against 106 us with two rounds in flight; the 512-matrix
// batch's one-segment plans 0.726 against 0.787 ms per pass.
__host__ __device__ inline bool blk_two_ahead(int S) { return S == 1; }

// Sum of the S partials of virtual row v in segment order, loads four at a
// time (S is 1, 4 or a multiple of 8).
__device__ __forceinline__ float blk_partials(const float* __restrict__ partial,
                                              long long part_stride, int S, long long v) {
  if (S == 1) return partial[v];
  float sum = 0.f;
  for (int s0 = 0; s0 < S; s0 += 4) {
    float g[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) g[u] = partial[(long long)(s0 + u) * part_stride + v];
    sum = s0 == 0 ? ((g[0] + g[1]) + g[2]) + g[3] : (((sum + g[0]) + g[1]) + g[2]) + g[3];
  }
  return sum;
}

// y of row r: ((partial_0 + partial_1) + ...) + partial_{S-1}; a row cut
// into pieces adds its pieces' sums in piece order.
__device__ __forceinline__ float blk_row_sum(const float* __restrict__ partial,
                                             long long part_stride, int S,
                                             const int* __restrict__ vrow, int r) {
  if (!vrow) return blk_partials(partial, part_stride, S, r);
  const int v0 = vrow[r], v1 = vrow[r + 1];
  float sum = blk_partials(partial, part_stride, S, v0);
  for (int v = v0 + 1; v < v1; ++v) sum += blk_partials(partial, part_stride, S, v);
  return sum;
}

}  // namespace loops_mi355x
