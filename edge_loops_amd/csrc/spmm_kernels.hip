// gfx950 SpMM kernels, C = A * B with B (cols x k) and C (rows x k) dense
// row-major fp32 (leading dimensions ldb, ldc).
//
//   thread_mapped  algorithms/spmm/thread_mapped.cuh:28-53: one row per
//                  work-item.  The reference rescans the row once per output
//                  column; here a work-item keeps 8 columns in registers and
//                  rescans once per 8 columns.  Each C(row, c) is still the
//                  sequential CSR-order fp32 sum.
//   group_mapped   (no reference SpMM): the block_mapped partition (a
//                  workgroup owns TPB rows) whose atoms are shared evenly by
//                  ranges of G lanes; the G lanes span the output columns
//                  (CPL columns per lane, one 16-byte load of a B row slice
//                  per lane when CPL = 4), so every B row is fetched as
//                  whole contiguous segments.  Rows split between ranges
//                  are summed in range order.
//   merge_path     (no reference SpMM): the same lane-per-column groups, but
//                  each group walks an equal share (IPU merge steps) of the
//                  merge path of rows and nonzeros, so one 200k-nonzero row
//                  is spread over many groups.  Rows crossing a share end are
//                  finished by a fixed-order fix-up (as for SpMV), per column.
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <climits>

#include <loops/schedule.hxx>
#include <loops/util/search.hxx>

#include "device_common.hpp"
#include "loops_internal.hpp"
#include "spmm_unit.hpp"

namespace loops_mi355x {

using loops::schedule::algorithms_t;

// ---------------------------------------------------------------------------
// thread_mapped
// ---------------------------------------------------------------------------
template <typename V>
__global__ void spmm_thread_mapped_kernel(
    loops::schedule::setup<algorithms_t::thread_mapped, 1, 1, int, int,
                           long long, long long> config,
    const int* __restrict__ indices, const V* __restrict__ values, int k,
    const float* __restrict__ B, long long ldb, float* __restrict__ C,
    long long ldc) {
  constexpr int CB = 8;
  for (auto row : config.tiles()) {
    for (int c0 = 0; c0 < k; c0 += CB) {
      float acc[CB];
#pragma unroll
      for (int j = 0; j < CB; ++j) acc[j] = 0.f;
      const int cn = k - c0 < CB ? k - c0 : CB;
      for (auto nz : config.atoms(row)) {
        const float v = value_traits<V>::load1(values + nz);
        const float* b = B + (long long)indices[nz] * ldb + c0;
#pragma unroll
        for (int j = 0; j < CB; ++j)
          if (j < cn) acc[j] += v * b[j];
      }
      float* c = C + (long long)row * ldc + c0;
#pragma unroll
      for (int j = 0; j < CB; ++j)
        if (j < cn) c[j] = acc[j];
    }
  }
}

// ---------------------------------------------------------------------------
// group_mapped: the block_mapped partition of the reference's schedule
// (group_mapped.hxx:146-213): a workgroup of TPB lanes owns TPB consecutive
// rows, and the group's atoms (its rows' nonzeros, numbered 0 .. aggregate-1
// by the exclusive scan of the row lengths) are shared out evenly.  The
// lanes form R = TPB / G ranges of G lanes spanning the columns; range q
// takes the contiguous atoms [q L, q L + L), L = ceil(aggregate / R), so a
// 200,000-nonzero row is spread over every range of its group instead of
// being walked by one.  Each range sums its atoms row by row (U B-row loads
// in flight, the next atoms' column ids loaded ahead) and writes the rows
// that end in its share; a row that began in an earlier share is finished
// by the range where it ends, adding the earlier ranges' partial sums in
// range order.  Every sum has an order fixed by the matrix, TPB and G.
// ---------------------------------------------------------------------------
template <int TPB, int G, int CPL, typename V>
__global__ __launch_bounds__(TPB) void spmm_group_mapped_kernel(
    const int* __restrict__ offsets, const int* __restrict__ indices,
    const V* __restrict__ values, int rows, int k, const float* __restrict__ B,
    long long ldb, float* __restrict__ C, long long ldc) {
  constexpr int R = TPB / G;
  constexpr int W = G * CPL;  // columns of one column block
  constexpr int U = 8;
  __shared__ int s_end[TPB + 1];        // s_end[i]: end of row r0 + i - 1, block-relative
  __shared__ int s_tail_row[R];         // row a range leaves open, or -1
  __shared__ float s_tail[R][W];        // its partial sums
  const int tid = threadIdx.x;
  const long long r0 = (long long)blockIdx.x * TPB;
  const int nrows = rows - r0 < TPB ? (int)(rows - r0) : TPB;
  const int base = offsets[r0];
  for (int i = tid; i <= nrows; i += TPB) s_end[i] = offsets[r0 + i] - base;
  __syncthreads();
  const int agg = s_end[nrows];
  const int L = (agg + R - 1) / R;
  const int q = tid / G, g = tid % G;
  const int a0 = q * L < agg ? q * L : agg;
  const int a1 = a0 + L < agg ? a0 + L : agg;
  // first block row whose end exceeds v (rows are 0 .. nrows-1)
  auto row_after = [&](int v) {
    int lo = 0, hi = nrows;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (s_end[mid + 1] > v) hi = mid;
      else lo = mid + 1;
    }
    return lo;
  };
  // rows ending in this share: end in (a0, a1]; the first range also owns
  // the rows ending at 0 (empty rows before the first nonzero)
  const int rlo = q == 0 ? 0 : row_after(a0);
  const int rhi = row_after(a1);
  const int* __restrict__ idx = indices + base;
  const V* __restrict__ val = values + base;

  for (int cb = 0; cb < k; cb += W) {
    const int c = cb + g * CPL;
    float acc[CPL];
#pragma unroll
    for (int j = 0; j < CPL; ++j) acc[j] = 0.f;
    int cr = rlo;          // current row
    int next_end = cr < nrows ? s_end[cr + 1] : INT_MAX;
    int head_row = -1;     // a finished row that began before a0
    float head[CPL];
#pragma unroll
    for (int j = 0; j < CPL; ++j) head[j] = 0.f;
    auto finish = [&](int r) {
      if (s_end[r] < a0) {
        head_row = r;
#pragma unroll
        for (int j = 0; j < CPL; ++j) head[j] = acc[j];
      } else if (c < k) {
        store_c<CPL>(C + (r0 + r) * ldc + c, c, k, acc);
      }
#pragma unroll
      for (int j = 0; j < CPL; ++j) acc[j] = 0.f;
    };
    auto walk = [&](auto whole) {
      constexpr bool kWhole = decltype(whole)::value;
      // kWhole: every lane of the wavefront has whole column slices and the
      // block holds nonzeros, so every load is issued unconditionally
      // (atoms past a1 read the block's first atom, their B rows are read
      // and not summed): with a load under a branch or a select on a loaded
      // value the compiler waited for each B row before the next
      auto atom = [&](int i) { return kWhole ? (i < a1 ? i : 0) : i; };
      int cj[U];
      float vj[U];
#pragma unroll
      for (int j = 0; j < U; ++j) {
        if constexpr (kWhole) {
          cj[j] = ld_idx(idx + atom(a0 + j));
          vj[j] = value_traits<V>::load1(val + atom(a0 + j));
        } else {
          cj[j] = a0 + j < a1 ? ld_idx(idx + a0 + j) : 0;
          vj[j] = a0 + j < a1 ? value_traits<V>::load1(val + a0 + j) : 0.f;
        }
      }
      for (int a = a0; a < a1; a += U) {
        const int n = a1 - a < U ? a1 - a : U;
        cols_t<CPL> bj[U];
#pragma unroll
        for (int j = 0; j < U; ++j) {
          if constexpr (kWhole) {
            const float* p = B + (long long)cj[j] * ldb + c;
            if constexpr (CPL == 4) {
              const f32x4 qv = *reinterpret_cast<const f32x4*>(p);
              bj[j].v[0] = qv.x; bj[j].v[1] = qv.y; bj[j].v[2] = qv.z; bj[j].v[3] = qv.w;
            } else {
              bj[j].v[0] = *p;
            }
          } else if (j < n) {
            bj[j] = load_b<CPL>(B, ldb, cj[j], c, k);
          }
        }
        float vcur[U];
#pragma unroll
        for (int j = 0; j < U; ++j) vcur[j] = vj[j];
        // the next atoms' column ids and values, in flight with the B rows
        const int an = a + U;
#pragma unroll
        for (int j = 0; j < U; ++j) {
          if constexpr (kWhole) {
            cj[j] = ld_idx(idx + atom(an + j));
            vj[j] = value_traits<V>::load1(val + atom(an + j));
          } else {
            cj[j] = an + j < a1 ? ld_idx(idx + an + j) : 0;
            vj[j] = an + j < a1 ? value_traits<V>::load1(val + an + j) : 0.f;
          }
        }
#pragma unroll
        for (int j = 0; j < U; ++j) {
          if (j < n) {
            while (next_end <= a + j) {
              finish(cr);
              ++cr;
              next_end = s_end[cr + 1];
            }
#pragma unroll
            for (int t = 0; t < CPL; ++t) acc[t] += vcur[j] * bj[j].v[t];
          }
        }
      }
    };
    if (agg > 0 && __all(c + CPL <= k)) walk(std::integral_constant<bool, true>{});
    else walk(std::integral_constant<bool, false>{});
    while (cr < rhi) {
      finish(cr);
      ++cr;
    }
    // the row left open: it has atoms in this share and ends in a later one
    const bool open = a1 > a0 && cr < nrows && s_end[cr] < a1 && s_end[cr + 1] > a1;
    if (g == 0) s_tail_row[q] = open ? cr : -1;
#pragma unroll
    for (int j = 0; j < CPL; ++j) s_tail[q][g * CPL + j] = acc[j];
    __syncthreads();
    if (head_row >= 0) {
      int q0 = q;
      while (q0 > 0 && s_tail_row[q0 - 1] == head_row) --q0;
      float v[CPL];
#pragma unroll
      for (int j = 0; j < CPL; ++j) v[j] = s_tail[q0][g * CPL + j];
      for (int p = q0 + 1; p < q; ++p) {
#pragma unroll
        for (int j = 0; j < CPL; ++j) v[j] = v[j] + s_tail[p][g * CPL + j];
      }
#pragma unroll
      for (int j = 0; j < CPL; ++j) v[j] = v[j] + head[j];
      if (c < k) store_c<CPL>(C + (r0 + head_row) * ldc + c, c, k, v);
    }
    __syncthreads();  // s_tail is reused by the next column block
  }
}

// One wavefront per 64 / L consecutive tiles, L lanes per tile
// (spmm_fixup_lanes): the lanes of a tile probe it; a row whose carries lie
// in one tile is finished by them, 4 columns per lane; the other rows by
// the wavefront, one after another.
__global__ __launch_bounds__(256) void spmm_fixup_kernel(const int* __restrict__ offsets,
                                                         int rows, int k, long long ti,
                                                         long long tiles, int L, spmm_ws ws,
                                                         float* __restrict__ C, long long ldc) {
  const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const long long t0 = wave * (64 / L);
  const int sub = lane & (L - 1);
  const long long t = t0 + lane / L;
  long long t_end = 0;
  const int r = t < tiles ? spmm_fixup_probe(offsets, rows, ti, t, ws, &t_end) : -1;
  const bool alone = r >= 0 && t_end == t + 1;
  if (alone) spmm_fixup_short_part(k, t, t_end, r, ws, C, ldc, sub, L);
  unsigned long long todo = __ballot(r >= 0 && !alone && sub == 0);
  while (todo) {
    const int src = __ffsll((long long)todo) - 1;
    todo &= todo - 1;
    spmm_fixup_row(k, t0 + src / L, __shfl(t_end, src, 64), __shfl(r, src, 64), ws, C, ldc);
  }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
namespace {
inline hipError_t record(hipEvent_t e, hipStream_t s) {
  return e ? hipEventRecord(e, s) : hipSuccess;
}

template <typename V>
hipError_t tm_spmm_t(const csr_args& A, int k, const float* B, long long ldb,
                     float* C, long long ldc, int tpb, hipEvent_t ev0,
                     hipEvent_t ev1, hipStream_t s) {
  hipError_t e = record(ev0, s);
  if (e != hipSuccess) return e;
  long long grid = ((long long)A.rows + tpb - 1) / tpb;
  if (grid > 0 && k > 0) {
    loops::schedule::setup<algorithms_t::thread_mapped, 1, 1, int, int,
                           long long, long long>
        config(const_cast<int*>(A.offsets), A.rows, A.nnz);
    spmm_thread_mapped_kernel<V><<<dim3((unsigned)grid), dim3(tpb), 0, s>>>(
        config, A.indices, static_cast<const V*>(A.values), k, B, ldb, C, ldc);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return record(ev1, s);
}

// Lanes per group and columns per lane for a given k and B/C alignment.
struct spmm_shape {
  int g;
  int cpl;
};
inline spmm_shape pick_shape(int k, bool vec4) {
  if (vec4 && k >= kSpmmVec4MinK) {
    int g = kSpmmMinG;
    while (g < 64 && g * 4 < k) g *= 2;
    return {g, 4};
  }
  int g = 1;
  while (g < 64 && g < k) g *= 2;
  return {g, 1};
}

template <int TPB, int G, int CPL, typename V>
hipError_t gm_spmm_t(const csr_args& A, int k, const float* B, long long ldb,
                     float* C, long long ldc, hipStream_t s) {
  const long long grid = ((long long)A.rows + TPB - 1) / TPB;
  spmm_group_mapped_kernel<TPB, G, CPL, V><<<dim3((unsigned)grid), dim3(TPB), 0, s>>>(
      A.offsets, A.indices, static_cast<const V*>(A.values), A.rows, k, B, ldb,
      C, ldc);
  return hipGetLastError();
}

template <int IPT, int CPL, typename V>
hipError_t mp_spmm_t(const csr_args& A, int k, const float* B, long long ldb,
                     float* C, long long ldc, int g, void* ws, hipStream_t s,
                     hipEvent_t ev1) {
  const long long total = (long long)A.rows + A.nnz;
  const long long ti = spmm_tile_items(g, IPT);
  const long long tiles = (total + ti - 1) / ti;
  spmm_ws w = carve_spmm_ws(ws, tiles, k);
  spmm_merge_path_kernel<IPT, CPL, V><<<dim3((unsigned)tiles), dim3(kSpmmTpb), 0, s>>>(
      A.offsets, A.indices, static_cast<const V*>(A.values), A.rows, A.nnz, k, B, ldb,
      C, ldc, g, w);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = record(ev1, s);
  if (e != hipSuccess) return e;
  if (tiles > 1) {
    const int L = spmm_fixup_lanes(k);
    const long long lanes = (tiles + 64 / L - 1) / (64 / L) * 64;
    spmm_fixup_kernel<<<dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, s>>>(
        A.offsets, A.rows, k, ti, tiles, L, w, C, ldc);
    e = hipGetLastError();
  }
  return e;
}

#define LOOPS_SPMM_SHAPES(X) \
  X(1, 1) X(2, 1) X(4, 1) X(8, 1) X(16, 1) X(32, 1) X(64, 1) \
  X(2, 4) X(4, 4) X(8, 4) X(16, 4) X(32, 4) X(64, 4)

template <typename V>
hipError_t gm_spmm_v(const csr_args& A, int k, const float* B, long long ldb,
                     float* C, long long ldc, spmm_shape sh, int tpb, hipStream_t s) {
#define X(G_, C_)                                                                         \
  if (sh.g == G_ && sh.cpl == C_) {                                                       \
    if (tpb == 256) return gm_spmm_t<256, G_, C_, V>(A, k, B, ldb, C, ldc, s);           \
    if (tpb == 512) return gm_spmm_t<512, G_, C_, V>(A, k, B, ldb, C, ldc, s);           \
    return gm_spmm_t<1024, G_, C_, V>(A, k, B, ldb, C, ldc, s);                          \
  }
  LOOPS_SPMM_SHAPES(X)
#undef X
  return hipErrorInvalidValue;
}

template <typename V>
hipError_t mp_spmm_v(const csr_args& A, int k, const float* B, long long ldb,
                     float* C, long long ldc, spmm_shape sh, int ipt, void* ws,
                     hipStream_t s, hipEvent_t ev1) {
#define X(I_, C_)                 \
  if (ipt == I_ && sh.cpl == C_) \
    return mp_spmm_t<I_, C_, V>(A, k, B, ldb, C, ldc, sh.g, ws, s, ev1);
  X(4, 1) X(8, 1) X(16, 1) X(32, 1) X(4, 4) X(8, 4) X(16, 4) X(32, 4)
#undef X
  return hipErrorInvalidValue;
}

// 16-byte B loads: B's rows 16-byte aligned.
inline bool vec4_ok(const float* B, long long ldb) {
  return (reinterpret_cast<uintptr_t>(B) % 16 == 0) && (ldb % 4 == 0);
}

// Lanes per range of the merge-path SpMM tile (at least kSpmmMinG).
spmm_shape tile_shape(int k, bool vec4) {
  spmm_shape sh = pick_shape(k, vec4);
  if (sh.g < kSpmmMinG) sh.g = kSpmmMinG;
  return sh;
}

inline int default_ipt(const spmm_shape& sh) { return spmm_default_ipt(sh.g); }

}  // namespace

hipError_t launch_tm_spmm(const csr_args& A, int k, const float* B, long long ldb,
                          float* C, long long ldc, int tpb, hipEvent_t ev0,
                          hipEvent_t ev1, hipStream_t s) {
  switch (A.vtype) {
    case LOOPS_VALUE_F32: return tm_spmm_t<float>(A, k, B, ldb, C, ldc, tpb, ev0, ev1, s);
    case LOOPS_VALUE_BF16: return tm_spmm_t<bf16_t>(A, k, B, ldb, C, ldc, tpb, ev0, ev1, s);
    case LOOPS_VALUE_F16: return tm_spmm_t<f16_t>(A, k, B, ldb, C, ldc, tpb, ev0, ev1, s);
    default: break;
  }
  return hipErrorInvalidValue;
}

bool gm_spmm_supported(int tpb) { return tpb == 256 || tpb == 512 || tpb == 1024; }

hipError_t launch_gm_spmm(const csr_args& A, int k, const float* B, long long ldb,
                          float* C, long long ldc, int tpb, hipEvent_t ev0,
                          hipEvent_t ev1, hipStream_t s) {
  hipError_t e = record(ev0, s);
  if (e != hipSuccess) return e;
  if (A.rows > 0 && k > 0) {
    spmm_shape sh = pick_shape(k, vec4_ok(B, ldb));
    switch (A.vtype) {
      case LOOPS_VALUE_F32: e = gm_spmm_v<float>(A, k, B, ldb, C, ldc, sh, tpb, s); break;
      case LOOPS_VALUE_BF16: e = gm_spmm_v<bf16_t>(A, k, B, ldb, C, ldc, sh, tpb, s); break;
      case LOOPS_VALUE_F16: e = gm_spmm_v<f16_t>(A, k, B, ldb, C, ldc, sh, tpb, s); break;
      default: e = hipErrorInvalidValue; break;
    }
    if (e != hipSuccess) return e;
  }
  return record(ev1, s);
}

bool mp_spmm_supported(int tpb, int ipt) {
  return tpb == 256 && (ipt == 0 || ipt == 4 || ipt == 8 || ipt == 16 || ipt == 32);
}

bool mp_spmm_fits(int k, const float* B, long long ldb, int ipt) {
  if (ipt == 0) return true;
  for (int j = 0, w = 0; j < k; j += w) {
    w = k > kSpmmChunk ? spmm_chunk_width(k, j) : k;
    if (!spmm_ipt_ok(ipt, tile_shape(w, vec4_ok(B + j, ldb)).g)) return false;
  }
  return true;
}

// Tiles of the launch: the most over both shapes (4 columns per lane when B
// rows are 16-byte aligned, else 1), so the workspace fits any B.
long long mp_spmm_tiles(const csr_args& A, int k, int ipt) {
  long long most = 0;
  for (int v4 = 0; v4 < 2; ++v4) {
    const spmm_shape sh = tile_shape(k, v4 != 0);
    const long long ti = spmm_tile_items(sh.g, ipt ? ipt : default_ipt(sh));
    const long long t = ((long long)A.rows + A.nnz + ti - 1) / ti;
    if (t > most) most = t;
  }
  return most;
}

size_t mp_spmm_ws_bytes(long long tiles, int k) {
  return tiles > 0 && k > 0 ? spmm_ws_bytes_impl(tiles, k) : 0;
}

size_t mp_spmm_ws_need(const csr_args& A, int k, int ipt) {
  if (k <= kSpmmChunk) return mp_spmm_ws_bytes(mp_spmm_tiles(A, k, ipt), k);
  size_t most = 0;
  int prev = 0;
  for (long long j = 0; j < k;) {
    const int w = spmm_chunk_width(k, j);
    if (w != prev) most = std::max(most, mp_spmm_ws_bytes(mp_spmm_tiles(A, w, ipt), w));
    prev = w;
    j += w;
  }
  return most;
}

hipError_t launch_mp_spmm(const csr_args& A, int k, const float* B, long long ldb,
                          float* C, long long ldc, int tpb, int ipt, void* ws,
                          hipEvent_t ev0, hipEvent_t ev1, hipStream_t s) {
  (void)tpb;
  hipError_t e = record(ev0, s);
  if (e != hipSuccess) return e;
  if ((long long)A.rows + A.nnz == 0 || k == 0 || A.rows == 0) return record(ev1, s);
  // column chunks (spmm_chunk_width), in order; the stop event after the
  // last chunk's main kernel
  for (int j = 0, w = 0; j < k; j += w) {
    w = k > kSpmmChunk ? spmm_chunk_width(k, j) : k;
    const bool last = j + w >= k;
    const float* Bj = B + j;
    float* Cj = C + j;
    const spmm_shape sh = tile_shape(w, vec4_ok(Bj, ldb));
    const int it = ipt ? ipt : default_ipt(sh);
    hipEvent_t stop = last ? ev1 : nullptr;
    switch (A.vtype) {
      case LOOPS_VALUE_F32: e = mp_spmm_v<float>(A, w, Bj, ldb, Cj, ldc, sh, it, ws, s, stop); break;
      case LOOPS_VALUE_BF16: e = mp_spmm_v<bf16_t>(A, w, Bj, ldb, Cj, ldc, sh, it, ws, s, stop); break;
      case LOOPS_VALUE_F16: e = mp_spmm_v<f16_t>(A, w, Bj, ldb, Cj, ldc, sh, it, ws, s, stop); break;
      default: e = hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace loops_mi355x
