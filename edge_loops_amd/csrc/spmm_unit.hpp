// SpMM building blocks shared by the single-matrix kernels and the batch
// kernel: column slices of B and C per lane, the merge-path SpMM tile and
// its fixed-order fix-up.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include <loops/util/search.hxx>

#include "device_common.hpp"
#include "loops_internal.hpp"

namespace loops_mi355x {

// ---------------------------------------------------------------------------
// Column slice of one lane: CPL consecutive columns starting at c.
// ---------------------------------------------------------------------------
template <int CPL>
struct cols_t {
  float v[CPL];
};

template <int CPL>
__device__ __forceinline__ cols_t<CPL> load_b(const float* __restrict__ B,
                                              long long ldb, int col, int c,
                                              int k) {
  cols_t<CPL> r;
  const float* p = B + (long long)col * ldb + c;
  if constexpr (CPL == 4) {
    if (c + 4 <= k) {
      f32x4 q = *reinterpret_cast<const f32x4*>(p);
      r.v[0] = q.x; r.v[1] = q.y; r.v[2] = q.z; r.v[3] = q.w;
      return r;
    }
  }
#pragma unroll
  for (int j = 0; j < CPL; ++j) r.v[j] = (c + j < k) ? p[j] : 0.f;
  return r;
}

template <int CPL>
__device__ __forceinline__ void store_c(float* __restrict__ p, int c, int k,
                                        const float* acc) {
  if constexpr (CPL == 4) {
    // rows of C (or of the fix-up slots, k floats each) need not be
    // 16-byte aligned
    if (c + 4 <= k && (reinterpret_cast<uintptr_t>(p) & 15) == 0) {
      f32x4 q;
      q.x = acc[0]; q.y = acc[1]; q.z = acc[2]; q.w = acc[3];
      *reinterpret_cast<f32x4*>(p) = q;
      return;
    }
  }
#pragma unroll
  for (int j = 0; j < CPL; ++j)
    if (c + j < k) p[j] = acc[j];
}

// ---------------------------------------------------------------------------
// Fix-up slots of the SpMM tiles: per tile the open row (int) and its carry
// and head partial sums (k fp32 each).
// ---------------------------------------------------------------------------
struct spmm_ws {
  int* carry_row;
  float* carry;  // [units][k]
  float* head;   // [units][k]
};

inline size_t spmm_ws_bytes_impl(long long units, int k) {
  size_t b = (size_t)units * sizeof(int);
  b = (b + 255) & ~size_t(255);
  b += 2 * (size_t)units * k * sizeof(float);
  return (b + 255) & ~size_t(255);
}

inline spmm_ws carve_spmm_ws(void* ws, long long units, int k) {
  char* p = static_cast<char*>(ws);
  size_t b = ((size_t)units * sizeof(int) + 255) & ~size_t(255);
  spmm_ws w;
  w.carry_row = reinterpret_cast<int*>(p);
  w.carry = reinterpret_cast<float*>(p + b);
  w.head = w.carry + (size_t)units * k;
  return w;
}

// ---------------------------------------------------------------------------
// Merge-path SpMM tile, C = A * B for k columns.
//
// A workgroup of TPB lanes takes one merge tile.  Its lanes form R = TPB / G
// ranges of G lanes (G = 2 .. 64, a power of two); the G lanes of a range
// span G * CPL consecutive columns of B and C, and the range walks IPT merge
// steps, so a tile covers TI = R * IPT steps.
//  1. Tile end points: read from a precomputed partition, or found by two
//     wavefronts with the 64-way cooperative search.
//  2. Row ends, column ids and values of the tile go to LDS (coalesced).
//  3. Each range finds its start and end in LDS, then issues the B-row loads
//     of up to 8 nonzeros at once (one round trip per 8 steps) and walks
//     those steps, storing each row it closes except the first, which it
//     holds back.
//  4. A segmented scan over the ranges (per column) carries partial sums
//     between ranges; the held-back rows are completed with it.
//  5. The row entering the tile open goes to head[tile], the row leaving it
//     open to carry[tile]; spmm_fixup_unit sums them in tile order.
// Column blocks of G * CPL columns are walked one after another when k is
// wider.  Every sum has an order fixed by the matrix, k, G, CPL and IPT.
// ---------------------------------------------------------------------------
constexpr int kSpmmTpb = 256;
constexpr int kSpmmMinG = 2;

template <int IPT, int CPL>
struct spmm_smem {
  // tiles of up to 16 steps per range for ranges of kSpmmMinG lanes or
  // more; 32 steps for ranges of 8 lanes or more (spmm_ipt_ok)
  static constexpr int TI_MAX = (kSpmmTpb / (IPT <= 16 ? kSpmmMinG : 8)) * IPT;
  static constexpr int NW = kSpmmTpb / 64;
  int rowend[TI_MAX + 1];
  int col[TI_MAX];
  float val[TI_MAX];
  int coord[5];
  float wv[NW][64 * CPL];
  int wf[NW][64];
};

// Merge steps per tile for lane-group size g.
__host__ __device__ constexpr long long spmm_tile_items(int g, int ipt) {
  return (long long)(kSpmmTpb / g) * ipt;
}

template <int IPT, int CPL, typename V>
__device__ __forceinline__ void spmm_tile(
    const int* __restrict__ offsets, const int* __restrict__ indices,
    const V* __restrict__ values, int rows, int nnz, int k,
    const float* __restrict__ B, long long ldb, float* __restrict__ C,
    long long ldc, int G, int tile, const loops_coord_t* __restrict__ coords,
    const spmm_ws& ws, spmm_smem<IPT, CPL>& sm) {
  constexpr int TPB = kSpmmTpb;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int TI = (TPB / G) * IPT;
  const long long total = (long long)rows + nnz;
  long long d0 = (long long)tile * TI;
  if (d0 > total) d0 = total;
  long long d1 = d0 + TI;
  if (d1 > total) d1 = total;

  // 1. tile end points
  if (coords != nullptr) {
    if (tid < 2) {
      const loops_coord_t c = coords[tile + tid];
      sm.coord[2 * tid] = (int)c.x;
      sm.coord[2 * tid + 1] = (int)c.y;
    }
  } else {
    if (wave == 0) {
      auto c = loops::search::wave_search(offsets, rows, nnz, d0);
      if (lane == 0) {
        sm.coord[0] = (int)c.x;
        sm.coord[1] = (int)c.y;
      }
    }
    if (wave == 1) {
      auto c = loops::search::wave_search(offsets, rows, nnz, d1);
      if (lane == 0) {
        sm.coord[2] = (int)c.x;
        sm.coord[3] = (int)c.y;
      }
    }
  }
  __syncthreads();
  const int xs = sm.coord[0], ys = sm.coord[1];
  const int xe = sm.coord[2], ye = sm.coord[3];
  const int tile_rows = xe - xs;
  const int tile_nnz = ye - ys;
  const int tile_items = (int)(d1 - d0);

  // 2. stage the tile: column ids and values, then row ends
  for (int i = tid; i < tile_nnz; i += TPB) {
    sm.col[i] = ld_idx(indices + ys + i);
    sm.val[i] = value_traits<V>::load1(values + ys + i);
  }
  for (int i = tid; i <= tile_rows; i += TPB) {
    int r = xs + 1 + i;
    if (r > rows) r = rows;
    sm.rowend[i] = offsets[r] - ys;
  }
  if (tid == 0) sm.coord[4] = offsets[xs] - ys;
  __syncthreads();

  // 3a. this range's start and end in the tile
  const int r = tid / G;
  const int g = tid & (G - 1);
  const int RW = 64 / G;         // ranges per wavefront
  const int rw = lane / G;       // this range's place in its wavefront
  auto lds_search = [&](int dd, int* xo, int* yo) {
    int x_min = dd - tile_nnz;
    x_min = x_min < 0 ? 0 : x_min;
    const int x_max = dd < tile_rows ? dd : tile_rows;
    int first = x_min, count = x_max - x_min;
    while (count > 0) {
      const int step = count >> 1;
      const int idx = first + step;
      if (sm.rowend[idx] <= dd - idx - 1) {
        first = idx + 1;
        count -= step + 1;
      } else {
        count = step;
      }
    }
    *xo = first;
    *yo = dd - first;
  };
  int dl = r * IPT;
  dl = dl < tile_items ? dl : tile_items;
  int dr = r * IPT + IPT;
  dr = dr < tile_items ? dr : tile_items;
  int xl, yl, xl2, yl2;
  lds_search(dl, &xl, &yl);
  lds_search(dr, &xl2, &yl2);
  const int cnt = yl2 - yl;  // nonzeros this range consumes (<= IPT)
  const bool open_left = tile_rows > 0 && sm.coord[4] < 0;
  const int xe_start = tile_rows > 0 ? sm.rowend[tile_rows - 1] : sm.coord[4];
  const bool open_right = xe < rows && xe_start < tile_nnz;

  for (int cb = 0; cb < k; cb += G * CPL) {
    const int c = cb + g * CPL;
    // 3b + 3c: the B rows of this range's nonzeros, 8 loads in flight
    // together, then the walk over its steps.
    float acc[CPL], first_acc[CPL];
    int x = xl, first_row = -1;
    auto load_walk = [&](auto whole) {
      constexpr bool kWhole = decltype(whole)::value;
      // B rows are loaded H steps at a time and walked before the next H
      // are loaded: the same sums in the same order as loading all IPT at
      // once, in a fixed register budget
      constexpr int H = IPT < 8 ? IPT : 8;
#pragma unroll
      for (int q = 0; q < CPL; ++q) acc[q] = first_acc[q] = 0.f;
      int row_end = sm.rowend[x];
      auto close_row = [&]() {
        if (first_row < 0) {
          first_row = x;
#pragma unroll
          for (int q = 0; q < CPL; ++q) first_acc[q] = acc[q];
        } else if (c < k) {
          store_c<CPL>(C + (long long)(xs + x) * ldc + c, c, k, acc);
        }
#pragma unroll
        for (int q = 0; q < CPL; ++q) acc[q] = 0.f;
        ++x;
        row_end = sm.rowend[x];
      };
      int yy = yl;
#pragma unroll
      for (int h = 0; h < IPT; h += H) {
        cols_t<CPL> bj[H];
        float vj[H];
#pragma unroll
        for (int j = 0; j < H; ++j) {
          if constexpr (kWhole) {
            // every lane's columns lie inside B's rows and the tile holds
            // nonzeros: a load per step whatever cnt is (steps past cnt
            // read the tile's first B row and are not summed), so no load
            // sits under a branch and none is waited for before the walk
            const int jj = h + j < cnt ? yl + h + j : 0;
            vj[j] = sm.val[jj];
            const float* p = B + (long long)sm.col[jj] * ldb + c;
            if constexpr (CPL == 4) {
              const f32x4 q = *reinterpret_cast<const f32x4*>(p);
              bj[j].v[0] = q.x; bj[j].v[1] = q.y; bj[j].v[2] = q.z; bj[j].v[3] = q.w;
            } else {
              bj[j].v[0] = *p;
            }
          } else if (h + j < cnt) {
            vj[j] = sm.val[yl + h + j];
            bj[j] = load_b<CPL>(B, ldb, sm.col[yl + h + j], c, k);
          }
        }
#pragma unroll
        for (int j = 0; j < H; ++j) {
          if (h + j < cnt) {
            while (row_end <= yy) close_row();
#pragma unroll
            for (int q = 0; q < CPL; ++q) acc[q] += vj[j] * bj[j].v[q];
            ++yy;
          }
        }
      }
      while (x < xl2) close_row();
    };
    // A wavefront whose lanes all have whole column slices (CPL columns
    // inside k) takes the branch-free loads: with a load under a branch the
    // compiler waited for each B row before issuing the next.
    if (tile_nnz > 0 && __all(c + CPL <= k))
      load_walk(std::integral_constant<bool, true>{});
    else
      load_walk(std::integral_constant<bool, false>{});

    // 4. segmented scan over the ranges, per column
    int f = first_row >= 0 ? 1 : 0;
    float v[CPL];
#pragma unroll
    for (int q = 0; q < CPL; ++q) v[q] = acc[q];
    for (int d = 1; d < RW; d <<= 1) {
      float vu[CPL];
#pragma unroll
      for (int q = 0; q < CPL; ++q) vu[q] = __shfl_up(v[q], d * G, 64);
      const int fu = __shfl_up(f, d * G, 64);
      if (rw >= d) {
        if (!f) {
#pragma unroll
          for (int q = 0; q < CPL; ++q) v[q] = vu[q] + v[q];
        }
        f |= fu;
      }
    }
    if (rw == RW - 1) {
#pragma unroll
      for (int q = 0; q < CPL; ++q) sm.wv[wave][g * CPL + q] = v[q];
      sm.wf[wave][g] = f;
    }
    __syncthreads();
    float pv[CPL];
#pragma unroll
    for (int q = 0; q < CPL; ++q) pv[q] = 0.f;
    for (int w = 0; w < wave; ++w) {
      const int wf = sm.wf[w][g];
#pragma unroll
      for (int q = 0; q < CPL; ++q) pv[q] = wf ? sm.wv[w][g * CPL + q] : pv[q] + sm.wv[w][g * CPL + q];
    }
    float ev[CPL];
#pragma unroll
    for (int q = 0; q < CPL; ++q) ev[q] = __shfl_up(v[q], G, 64);
    const int ef = __shfl_up(f, G, 64);
    if (first_row >= 0 && c < k) {
      float val[CPL];
#pragma unroll
      for (int q = 0; q < CPL; ++q)
        val[q] = (rw == 0 ? pv[q] : (ef ? ev[q] : pv[q] + ev[q])) + first_acc[q];
      if (first_row == 0 && open_left) {
        store_c<CPL>(ws.head + (long long)tile * k + c, c, k, val);
      } else {
        store_c<CPL>(C + (long long)(xs + first_row) * ldc + c, c, k, val);
      }
    }
    // 5. the row leaving the tile open: the last range's inclusive value
    if (r == TPB / G - 1) {
      float tot[CPL];
#pragma unroll
      for (int q = 0; q < CPL; ++q) tot[q] = f ? v[q] : pv[q] + v[q];
      if (c < k) store_c<CPL>(ws.carry + (long long)tile * k + c, c, k, tot);
      if (cb == 0 && g == 0) ws.carry_row[tile] = open_right ? xe : -1;
    }
    __syncthreads();  // sm.wv / sm.wf are reused by the next column block
  }
}

template <int IPT, int CPL, typename V>
__global__ __launch_bounds__(kSpmmTpb) void spmm_merge_path_kernel(
    const int* __restrict__ offsets, const int* __restrict__ indices,
    const V* __restrict__ values, int rows, int nnz, int k,
    const float* __restrict__ B, long long ldb, float* __restrict__ C,
    long long ldc, int G, spmm_ws ws) {
  __shared__ spmm_smem<IPT, CPL> sm;
  spmm_tile<IPT, CPL, V>(offsets, indices, values, rows, nnz, k, B, ldb, C, ldc, G,
                         blockIdx.x, nullptr, ws, sm);
}

// Fix-up of rows that cross tiles.  A row r open across tiles t .. t_end
// (t = the tile where it starts, fr / ti with fr = offsets[r] + r its first
// merge step) is finished as carry[t] + carry[t+1] + ... + head[t_end] per
// column, in tile order.
//
// spmm_fixup_probe: for tile t, the row to finish there (or -1) and t_end.
__device__ __forceinline__ int spmm_fixup_probe(const int* __restrict__ offsets, int rows,
                                                long long ti, long long t,
                                                const spmm_ws& ws, long long* t_end) {
  const int r = ws.carry_row[t];
  if (r < 0 || r >= rows) return -1;
  const long long fr = (long long)offsets[r] + r;
  if (fr / ti != t) return -1;
  *t_end = ((long long)offsets[r + 1] + r) / ti;
  return r;
}

// spmm_fixup_short_part: a row whose carries lie in one tile (t_end == t +
// 1), finished by L lanes (lane sub takes columns 4 sub, 4 (sub + L), ...
// where the slots allow 16-byte loads, else columns sub, sub + L, ...):
// C = (0 + carry[t]) + head[t_end] per column, the value spmm_fixup_row
// gives it (its other partitions add +0).  spmm_fixup_short: one lane.
__device__ __forceinline__ void spmm_fixup_short_part(int k, long long t, long long t_end,
                                                      int r, const spmm_ws& ws,
                                                      float* __restrict__ C, long long ldc,
                                                      int sub, int L) {
  const float* __restrict__ cr = ws.carry + t * k;
  const float* __restrict__ hr = ws.head + t_end * k;
  float* __restrict__ out = C + (long long)r * ldc;
  if (((reinterpret_cast<uintptr_t>(cr) | reinterpret_cast<uintptr_t>(hr) |
        (uintptr_t)k) & 3) == 0 &&
      ((reinterpret_cast<uintptr_t>(cr) | reinterpret_cast<uintptr_t>(hr)) & 15) == 0) {
    for (int c = 4 * sub; c < k; c += 4 * L) {
      const f32x4 a = *reinterpret_cast<const f32x4*>(cr + c);
      const f32x4 h = *reinterpret_cast<const f32x4*>(hr + c);
      out[c] = (0.f + a.x) + h.x;
      out[c + 1] = (0.f + a.y) + h.y;
      out[c + 2] = (0.f + a.z) + h.z;
      out[c + 3] = (0.f + a.w) + h.w;
    }
    return;
  }
  for (int c = sub; c < k; c += L) out[c] = (0.f + cr[c]) + hr[c];
}
__device__ __forceinline__ void spmm_fixup_short(int k, long long t, long long t_end, int r,
                                                 const spmm_ws& ws, float* __restrict__ C,
                                                 long long ldc) {
  spmm_fixup_short_part(k, t, t_end, r, ws, C, ldc, 0, 1);
}

// Lanes per tile of the fix-up kernel: one per 4 columns, a power of two,
// at most 16.
__host__ __device__ inline int spmm_fixup_lanes(int k) {
  int L = 1;
  while (L < 16 && 4 * L < k) L <<= 1;
  return L;
}

// spmm_fixup_row: the calling wavefront splits its 64 lanes into kp column
// lanes x 64 / kp tile partitions; partition p sums carries t + p,
// t + p + 64 / kp, ... < t_end for its columns, an xor butterfly adds the
// partitions, and head[t_end] is added last.  When k is a multiple of 4 a
// lane takes 4 columns (kp = k / 4 rounded up to a power of two, at most
// 64), so a 64-column row has 4 partitions instead of 1; else 1 column (kp
// = k rounded up, at most 64).  The order is fixed by k alone (16-byte
// loads where the slots allow, the same sums either way).  A partition
// loads eight carries ahead of its additions (a row of 200,000 nonzeros
// spans ~400 tiles of 512 steps), then adds them in order.
__device__ __forceinline__ void spmm_fixup_row(int k, long long t, long long t_end, int r,
                                               const spmm_ws& ws, float* __restrict__ C,
                                               long long ldc) {
  const int lane = threadIdx.x & 63;
  if ((k & 3) == 0) {
    const int kq = k >> 2;  // groups of 4 columns
    int kp = 1;
    while (kp < kq && kp < 64) kp <<= 1;
    const int parts = 64 / kp;
    const int cl = lane & (kp - 1), part = lane / kp;
    const bool vec = ((reinterpret_cast<uintptr_t>(ws.carry) |
                       reinterpret_cast<uintptr_t>(ws.head)) & 15) == 0;
    auto ld4 = [&](const float* p) {
      if (vec) return *reinterpret_cast<const f32x4*>(p);
      f32x4 q;
      q.x = p[0]; q.y = p[1]; q.z = p[2]; q.w = p[3];
      return q;
    };
    for (int q0 = 0; q0 < kq; q0 += kp) {
      const int c = 4 * (q0 + cl);
      f32x4 s = {0.f, 0.f, 0.f, 0.f};
      if (c < k) {
        long long v = t + part;
        for (; v + 7LL * parts < t_end; v += 8LL * parts) {
          f32x4 g[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) g[u] = ld4(ws.carry + (v + (long long)u * parts) * k + c);
#pragma unroll
          for (int u = 0; u < 8; ++u) s += g[u];
        }
        for (; v < t_end; v += parts) s += ld4(ws.carry + v * k + c);
      }
      for (int m = kp; m < 64; m <<= 1) {
        s.x += __shfl_xor(s.x, m, 64);
        s.y += __shfl_xor(s.y, m, 64);
        s.z += __shfl_xor(s.z, m, 64);
        s.w += __shfl_xor(s.w, m, 64);
      }
      if (part == 0 && c < k) {
        const f32x4 h = ld4(ws.head + t_end * k + c);
        const float out[4] = {s.x + h.x, s.y + h.y, s.z + h.z, s.w + h.w};
        store_c<4>(C + (long long)r * ldc + c, c, k, out);
      }
    }
    return;
  }
  int kp = 1;
  while (kp < k && kp < 64) kp <<= 1;
  const int parts = 64 / kp;
  const int cl = lane & (kp - 1), part = lane / kp;
  for (int c0 = 0; c0 < k; c0 += kp) {
    const int c = c0 + cl;
    float s = 0.f;
    if (c < k) {
      long long v = t + part;
      for (; v + 7LL * parts < t_end; v += 8LL * parts) {
        float g[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) g[u] = ws.carry[(v + (long long)u * parts) * k + c];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += g[u];
      }
      for (; v < t_end; v += parts) s += ws.carry[v * k + c];
    }
    for (int m = kp; m < 64; m <<= 1) s += __shfl_xor(s, m, 64);
    if (part == 0 && c < k) C[(long long)r * ldc + c] = s + ws.head[t_end * k + c];
  }
}

}  // namespace loops_mi355x
