// Column-panel SpMV plan: keep each XCD's share of x resident in its L2.
//
// A merge-path SpMV over a CSR matrix with columns spread over the whole of
// x misses the 4 MB per-XCD L2 on about half of its x reads once x exceeds a
// few MB, and every miss moves a 128-byte line.  The plan splits A by column
// into P panels of cols/P columns (columns are sorted within each row, so
// each row splits into P contiguous pieces) and stores each panel as its own
// CSR (all rows, that panel's nonzeros).  One launch runs every panel's merge
// tiles with tile b of the launch taking panel b mod P: workgroups are handed
// to the 8 XCDs round-robin, so panel p's tiles run on XCDs p, p+P, ... and
// each XCD's L2 only ever holds the x slice of one panel.  Each panel writes
// its own partial y (rows crossing its tiles are finished by the usual
// fix-up), and a last kernel adds the P partials of every row in panel
// order.  With P = 1 the plan is just the matrix's precomputed tile
// coordinates and fix-up workspace.
//
// Build (once per matrix, loops_spmv_plan_create): split counts per row and
// panel, an exclusive scan per panel (three passes), a scatter of column ids
// and values into the panel arrays, and the panels' tile coordinates.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <new>
#include <vector>

#include "device_common.hpp"
#include "loops_internal.hpp"
#include "merge_tile.hpp"

namespace loops_mi355x {

// ---------------------------------------------------------------------------
// Build kernels
// ---------------------------------------------------------------------------

// cnt[p * rows + r] = number of nonzeros of row r in columns [bound[p],
// bound[p+1]).
__global__ void panel_split_kernel(const int* __restrict__ offsets,
                                   const int* __restrict__ indices, int rows,
                                   int P, panel_bounds_t bound,
                                   int* __restrict__ cnt) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const int s = offsets[r], e = offsets[r + 1];
  int prev = s;
  for (int p = 0; p < P; ++p) {
    int hi;
    if (p == P - 1) {
      hi = e;
    } else {
      int lo = prev, h = e;  // first position with column >= bound[p+1]
      while (lo < h) {
        int mid = (lo + h) >> 1;
        if (indices[mid] < bound.c[p + 1]) lo = mid + 1;
        else h = mid;
      }
      hi = lo;
    }
    cnt[(long long)p * rows + r] = hi - prev;
    prev = hi;
  }
}

constexpr int kScanItems = 2048;  // per 256-lane block, 8 per lane

// bsum[p * nb + b] = sum of cnt[p][b*2048 .. b*2048 + 2047]
__global__ __launch_bounds__(256) void scan_block_sums_kernel(
    const int* __restrict__ cnt, int rows, int nb, int* __restrict__ bsum) {
  const int p = blockIdx.y, b = blockIdx.x;
  const int* c = cnt + (long long)p * rows;
  int s = 0;
  for (int i = 0; i < 8; ++i) {
    long long j = (long long)b * kScanItems + i * 256 + threadIdx.x;
    if (j < rows) s += c[j];
  }
  __shared__ int red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[(long long)p * nb + b] = red[0];
}

// In place: bsum[p][b] = exclusive prefix over b; total[p] = panel nnz.
__global__ __launch_bounds__(1024) void scan_block_prefix_kernel(
    int* __restrict__ bsum, int nb, int* __restrict__ total) {
  const int p = blockIdx.x;
  int* v = bsum + (long long)p * nb;
  __shared__ int buf[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 1024) {
    const int i = base + threadIdx.x;
    const int x = i < nb ? v[i] : 0;
    buf[threadIdx.x] = x;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
      int t = threadIdx.x >= d ? buf[threadIdx.x - d] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) v[i] = carry + buf[threadIdx.x] - x;
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) total[p] = carry;
}

// poff[p][j] = bsum[p][b] + exclusive prefix of cnt[p] inside block b;
// poff[p][rows] = total.
__global__ __launch_bounds__(256) void scan_write_kernel(
    const int* __restrict__ cnt, int rows, int nb, const int* __restrict__ bsum,
    const int* __restrict__ total, panel_ptrs_t poff) {
  const int p = blockIdx.y, b = blockIdx.x;
  const int* c = cnt + (long long)p * rows;
  int* out = poff.p[p];
  const long long j0 = (long long)b * kScanItems + threadIdx.x * 8;
  int v[8];
  int s = 0;
  for (int i = 0; i < 8; ++i) {
    v[i] = j0 + i < rows ? c[j0 + i] : 0;
    s += v[i];
  }
  __shared__ int buf[256];
  buf[threadIdx.x] = s;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {
    int t = threadIdx.x >= d ? buf[threadIdx.x - d] : 0;
    __syncthreads();
    buf[threadIdx.x] += t;
    __syncthreads();
  }
  int run = bsum[(long long)p * nb + b] + buf[threadIdx.x] - s;
  for (int i = 0; i < 8; ++i) {
    if (j0 + i < rows) out[j0 + i] = run;
    run += v[i];
  }
  if (b == 0 && threadIdx.x == 0) out[rows] = total[p];
}

// Exclusive scan of n ints (n >= 1): out[i] = cnt[0] + ... + cnt[i-1],
// out[n] = total.  scratch: ((n + 2047) / 2048 + 1) ints + one int.
size_t exclusive_scan_scratch_bytes(long long n) {
  return ((size_t)((n + kScanItems - 1) / kScanItems) + 2) * sizeof(int);
}

hipError_t launch_exclusive_scan(const int* cnt, long long n, int* out, void* scratch,
                                 hipStream_t s) {
  const int nb = (int)((n + kScanItems - 1) / kScanItems);
  int* bsum = static_cast<int*>(scratch);
  int* total = bsum + nb + 1;
  panel_ptrs_t po{};
  po.p[0] = out;
  scan_block_sums_kernel<<<dim3(nb, 1), 256, 0, s>>>(cnt, (int)n, nb, bsum);
  scan_block_prefix_kernel<<<1, 1024, 0, s>>>(bsum, nb, total);
  scan_write_kernel<<<dim3(nb, 1), 256, 0, s>>>(cnt, (int)n, nb, bsum, total, po);
  return hipGetLastError();
}

// Copies every nonzero to its panel: row by binary search over offsets,
// panel by column, position from the panel offsets and the row's counts.
template <typename V>
__global__ void panel_scatter_kernel(const int* __restrict__ offsets,
                                     const int* __restrict__ indices,
                                     const V* __restrict__ values, int rows,
                                     int nnz, int P, panel_bounds_t bound,
                                     const int* __restrict__ cnt,
                                     panel_ptrs_t poff, panel_ptrs_t pidx,
                                     panel_vptrs_t pval) {
  for (long long nz = (long long)blockIdx.x * blockDim.x + threadIdx.x; nz < nnz;
       nz += (long long)gridDim.x * blockDim.x) {
    int lo = 0, hi = rows;  // row r: offsets[r] <= nz < offsets[r+1]
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (offsets[mid] <= nz) lo = mid;
      else hi = mid;
    }
    const int r = lo;
    const int col = indices[nz];
    int p = 0;
    while (p + 1 < P && col >= bound.c[p + 1]) ++p;
    long long start = offsets[r];
    for (int q = 0; q < p; ++q) start += cnt[(long long)q * rows + r];
    const long long dst = (long long)poff.p[p][r] + (nz - start);
    pidx.p[p][dst] = col;
    static_cast<V*>(pval.p[p])[dst] = values[nz];
  }
}

// ---------------------------------------------------------------------------
// Execute kernels
// ---------------------------------------------------------------------------
template <int TPB, int IPT, typename V>
__global__ __launch_bounds__(TPB) void spmv_panels_kernel(panel_set_t ps,
                                                          const float* __restrict__ x,
                                                          int rows, int P) {
  __shared__ mp_smem<TPB, IPT> sm;
  const int p = (int)(blockIdx.x % (unsigned)P);
  const int t = (int)(blockIdx.x / (unsigned)P);
  if (t >= ps.tiles[p]) return;  // workgroup-uniform
  mp_tile<TPB, IPT, V, true>(ps.offsets[p], ps.indices[p],
                             static_cast<const V*>(ps.values[p]), x, ps.y[p], rows,
                             ps.nnz[p], t, ps.coords[p], ps.ws[p], sm);
}

// Fix-up of every panel: work-item u stands for tile u % tmax of panel
// u / tmax (same rules as spmv_fixup_kernel).
__global__ __launch_bounds__(256) void spmv_panels_fixup_kernel(panel_set_t ps,
                                                                int P, int tmax) {
  const long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int p = (int)(u / tmax);
  const long long t = u % tmax;
  long long t1 = -1;
  int r = -1;
  if (p < P && t < ps.tiles[p]) {
    t1 = ps.ws[p].seg_end[t];
    if (t1 >= 0) r = ps.ws[p].carry_row[t];
  }
  if (p >= P) p = P - 1;
  if (t1 >= 0 && segment_is_short(t, t1))
    ps.y[p][r] = short_segment_sum(ps.ws[p].carry_val, ps.ws[p].head_val, t, t1);
  unsigned long long todo = __ballot(t1 >= 0 && !segment_is_short(t, t1));
  while (todo) {
    const int src = __ffsll((long long)todo) - 1;
    todo &= todo - 1;
    const int pp = __shfl(p, src, 64);
    const long long a = __shfl(t, src, 64);
    const long long b = __shfl(t1, src, 64);
    const int rr = __shfl(r, src, 64);
    const float v = wave_segment_sum(ps.ws[pp].carry_val, ps.ws[pp].head_val, a, b);
    if (lane == src) ps.y[pp][rr] = v;
  }
}

// y[r] = ((y_0[r] + y_1[r]) + y_2[r]) + ...
__global__ __launch_bounds__(256) void spmv_panels_combine_kernel(panel_set_t ps,
                                                                  int P, int rows,
                                                                  float* __restrict__ y) {
  for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < rows;
       r += (long long)gridDim.x * blockDim.x) {
    float s = ps.y[0][r];
    for (int p = 1; p < P; ++p) s += ps.y[p][r];
    y[r] = s;
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

template <int TPB, int IPT>
hipError_t panels_launch(const panel_set_t& ps, loops_value_type_t vt,
                         const float* x, int rows, int P, int tmax,
                         hipStream_t s) {
  const unsigned grid = (unsigned)((long long)P * tmax);
  switch (vt) {
    case LOOPS_VALUE_F32:
      spmv_panels_kernel<TPB, IPT, float><<<grid, TPB, 0, s>>>(ps, x, rows, P);
      break;
    case LOOPS_VALUE_BF16:
      spmv_panels_kernel<TPB, IPT, bf16_t><<<grid, TPB, 0, s>>>(ps, x, rows, P);
      break;
    case LOOPS_VALUE_F16:
      spmv_panels_kernel<TPB, IPT, f16_t><<<grid, TPB, 0, s>>>(ps, x, rows, P);
      break;
    default:  // storage types of the typed entry points never reach a plan
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

bool panels_supported(int tpb, int ipt) {
  return (tpb == 256 && (ipt == 4 || ipt == 8 || ipt == 16)) ||
         (tpb == 512 && (ipt == 4 || ipt == 8));
}

hipError_t plan_build(const csr_args& A, int P, int tpb, int ipt, plan_state* st,
                      hipStream_t s) {
  st->P = P;
  st->rows = A.rows;
  st->vtype = A.vtype;
  st->tpb = tpb;
  st->ipt = ipt;
  const long long TI = (long long)tpb * ipt;
  const int vbytes = A.vtype == LOOPS_VALUE_F32 ? 4 : 2;
  panel_bounds_t bound{};
  for (int p = 0; p <= P; ++p) bound.c[p] = (int)(((long long)A.cols * p) / P);
  st->bound = bound;

  hipError_t e = hipSuccess;
  std::vector<long long> pnnz(P, 0);
  if (P == 1) {
    pnnz[0] = A.nnz;
  } else {
    // counts and per-panel offsets in a scratch + the plan's own memory
    const int nb = (int)((A.rows + kScanItems - 1) / kScanItems);
    size_t scratch = align256((size_t)P * A.rows * sizeof(int)) +
                     align256((size_t)P * (nb > 0 ? nb : 1) * sizeof(int)) +
                     align256((size_t)P * sizeof(int));
    char* sc = nullptr;
    if ((e = hipMalloc(&sc, scratch)) != hipSuccess) return e;
    int* cnt = reinterpret_cast<int*>(sc);
    int* bsum = reinterpret_cast<int*>(sc + align256((size_t)P * A.rows * sizeof(int)));
    int* total = reinterpret_cast<int*>(sc + align256((size_t)P * A.rows * sizeof(int)) +
                                        align256((size_t)P * (nb > 0 ? nb : 1) * sizeof(int)));
    // offsets of all panels
    size_t off_bytes = align256((size_t)(A.rows + 1) * sizeof(int));
    if ((e = hipMalloc(&st->offsets_mem, off_bytes * P)) != hipSuccess) {
      (void)hipFree(sc);
      return e;
    }
    panel_ptrs_t poff{};
    for (int p = 0; p < P; ++p)
      poff.p[p] = reinterpret_cast<int*>(static_cast<char*>(st->offsets_mem) + off_bytes * p);
    if (A.rows > 0) {
      panel_split_kernel<<<(unsigned)((A.rows + 255) / 256), 256, 0, s>>>(
          A.offsets, A.indices, A.rows, P, bound, cnt);
      scan_block_sums_kernel<<<dim3(nb, P), 256, 0, s>>>(cnt, A.rows, nb, bsum);
      scan_block_prefix_kernel<<<P, 1024, 0, s>>>(bsum, nb, total);
      scan_write_kernel<<<dim3(nb, P), 256, 0, s>>>(cnt, A.rows, nb, bsum, total, poff);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    std::vector<int> tot(P, 0);
    if (A.rows > 0) {
      if ((e = hipMemcpyAsync(tot.data(), total, P * sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess)
        return e;
      if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    }
    for (int p = 0; p < P; ++p) pnnz[p] = tot[p];
    // panel column ids and values, each panel 256-byte aligned
    size_t ib = 0, vb = 0;
    std::vector<size_t> ioff(P), voff(P);
    for (int p = 0; p < P; ++p) {
      ioff[p] = ib;
      voff[p] = vb;
      ib += align256((size_t)pnnz[p] * sizeof(int));
      vb += align256((size_t)pnnz[p] * vbytes);
    }
    if ((e = hipMalloc(&st->data_mem, ib + vb + 256)) != hipSuccess) return e;
    panel_ptrs_t pidx{};
    panel_vptrs_t pval{};
    for (int p = 0; p < P; ++p) {
      pidx.p[p] = reinterpret_cast<int*>(static_cast<char*>(st->data_mem) + ioff[p]);
      pval.p[p] = static_cast<char*>(st->data_mem) + ib + voff[p];
    }
    if (A.nnz > 0) {
      long long grid = ((long long)A.nnz + 255) / 256;
      if (grid > 65536) grid = 65536;
      switch (A.vtype) {
        case LOOPS_VALUE_F32:
          panel_scatter_kernel<float><<<(unsigned)grid, 256, 0, s>>>(
              A.offsets, A.indices, static_cast<const float*>(A.values), A.rows, A.nnz,
              P, bound, cnt, poff, pidx, pval);
          break;
        case LOOPS_VALUE_BF16:
          panel_scatter_kernel<bf16_t><<<(unsigned)grid, 256, 0, s>>>(
              A.offsets, A.indices, static_cast<const bf16_t*>(A.values), A.rows,
              A.nnz, P, bound, cnt, poff, pidx, pval);
          break;
        case LOOPS_VALUE_F16:
          panel_scatter_kernel<f16_t><<<(unsigned)grid, 256, 0, s>>>(
              A.offsets, A.indices, static_cast<const f16_t*>(A.values), A.rows,
              A.nnz, P, bound, cnt, poff, pidx, pval);
          break;
        default:
          break;  // check_csr admits fp32 / bf16 / fp16 only
      }
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    (void)hipFree(sc);
    for (int p = 0; p < P; ++p) {
      st->ps.offsets[p] = poff.p[p];
      st->ps.indices[p] = pidx.p[p];
      st->ps.values[p] = pval.p[p];
    }
  }
  if (P == 1) {
    st->ps.offsets[0] = A.offsets;
    st->ps.indices[0] = A.indices;
    st->ps.values[0] = A.values;
  }
  // tile coordinates, fix-up slots and partial outputs
  long long tmax = 0;
  size_t coord_bytes = 0, ws_bytes = 0;
  for (int p = 0; p < P; ++p) {
    const long long T = ((long long)A.rows + pnnz[p] + TI - 1) / TI;
    st->ps.nnz[p] = (int)pnnz[p];
    st->ps.tiles[p] = (int)T;
    if (T > tmax) tmax = T;
    coord_bytes += align256((size_t)(T + 1) * sizeof(loops_coord_t));
    ws_bytes += T > 0 ? fixup_ws_bytes(T) : 0;
  }
  st->tmax = (int)tmax;
  const size_t part_bytes = P > 1 ? align256((size_t)A.rows * sizeof(float)) * P : 0;
  if ((e = hipMalloc(&st->aux_mem, coord_bytes + ws_bytes + part_bytes + 256)) != hipSuccess)
    return e;
  char* q = static_cast<char*>(st->aux_mem);
  for (int p = 0; p < P; ++p) {
    const long long T = st->ps.tiles[p];
    loops_coord_t* c = reinterpret_cast<loops_coord_t*>(q);
    st->ps.coords[p] = c;
    q += align256((size_t)(T + 1) * sizeof(loops_coord_t));
    if ((e = launch_mp_partition(st->ps.offsets[p], A.rows, (int)pnnz[p], TI, T + 1, c, s)) !=
        hipSuccess)
      return e;
  }
  for (int p = 0; p < P; ++p) {
    const long long T = st->ps.tiles[p];
    st->ps.ws[p] = carve_fixup_ws(q, T);
    q += T > 0 ? fixup_ws_bytes(T) : 0;
  }
  for (int p = 0; p < P; ++p) {
    st->ps.y[p] = P > 1 ? reinterpret_cast<float*>(q) : nullptr;
    if (P > 1) q += align256((size_t)A.rows * sizeof(float));
  }
  st->device_bytes = coord_bytes + ws_bytes + part_bytes +
                     (P > 1 ? align256((size_t)(A.rows + 1) * sizeof(int)) * P : 0);
  for (int p = 0; p < P && P > 1; ++p)
    st->device_bytes += align256((size_t)pnnz[p] * sizeof(int)) + align256((size_t)pnnz[p] * vbytes);
  return hipStreamSynchronize(s);
}

hipError_t plan_execute(plan_state* st, const float* x, float* y, hipEvent_t ev0,
                        hipEvent_t ev1, hipStream_t s) {
  hipError_t e = ev0 ? hipEventRecord(ev0, s) : hipSuccess;
  if (e != hipSuccess) return e;
  const long long TI = (long long)st->tpb * st->ipt;
  if (st->P == 1) {
    csr_args A;
    A.offsets = st->ps.offsets[0];
    A.indices = st->ps.indices[0];
    A.values = st->ps.values[0];
    A.vtype = st->vtype;
    A.rows = st->rows;
    A.cols = 0;
    A.nnz = st->ps.nnz[0];
    A.aligned = st->aligned;
    return launch_mp_spmv(A, x, y, st->tpb, st->ipt, st->ps.coords[0], st->ps.ws[0],
                          nullptr, ev1, s);
  }
  if (st->tmax > 0) {
#define X(T, I)                                                                     \
  if (st->tpb == T && st->ipt == I)                                                 \
    e = panels_launch<T, I>(st->ps, st->vtype, x, st->rows, st->P, st->tmax, s); \
  else
    X(256, 4) X(256, 8) X(256, 16) X(512, 4) X(512, 8) e = hipErrorInvalidValue;
#undef X
    if (e != hipSuccess) return e;
  }
  if (ev1 && (e = hipEventRecord(ev1, s)) != hipSuccess) return e;
  (void)TI;
  const long long units = (long long)st->P * st->tmax;
  if (units > 0) {
    spmv_panels_fixup_kernel<<<(unsigned)((units + 255) / 256), 256, 0, s>>>(st->ps, st->P,
                                                                              st->tmax);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (st->rows > 0) {
    long long grid = ((long long)st->rows + 255) / 256;
    if (grid > 8192) grid = 8192;
    spmv_panels_combine_kernel<<<(unsigned)grid, 256, 0, s>>>(st->ps, st->P, st->rows, y);
    e = hipGetLastError();
  }
  return e;
}

void plan_release(plan_state* st) {
  if (st->offsets_mem) (void)hipFree(st->offsets_mem);
  if (st->data_mem) (void)hipFree(st->data_mem);
  if (st->aux_mem) (void)hipFree(st->aux_mem);
  st->offsets_mem = st->data_mem = st->aux_mem = nullptr;
}

// ---------------------------------------------------------------------------
// CSR structure check (loops_csr_check, and the first step of every plan
// build): offsets from 0 to nnz without decreasing, column ids in
// [0, cols), and ascending inside each row.  Work-item i checks row i and
// entry i; a descent from entry i to entry i + 1 is only allowed where
// i + 1 starts a row, found by a lower_bound over offsets[0 .. rows].  Every
// read is inside offsets[0 .. rows] and indices[0 .. nnz), whatever the
// values are.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void csr_check_kernel(const int* __restrict__ offsets,
                                                        const int* __restrict__ indices,
                                                        int rows, int nnz, int cols,
                                                        unsigned* __restrict__ flags) {
  unsigned f = 0;
  const long long n = rows > nnz ? rows : nnz;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    if (i < rows) {
      if (offsets[i] > offsets[i + 1]) f |= LOOPS_CSR_BAD_OFFSETS;
      if (i == 0 && offsets[0] != 0) f |= LOOPS_CSR_BAD_OFFSETS;
      if (i == rows - 1 && offsets[rows] != nnz) f |= LOOPS_CSR_BAD_OFFSETS;
    }
    if (i < nnz) {
      const int c = indices[i];
      if (c < 0 || c >= cols) f |= LOOPS_CSR_COLUMN_RANGE;
      if (i + 1 < nnz && c > indices[i + 1]) {
        const long long next = i + 1;
        // first r with offsets[r] >= next (64-bit: lo + hi exceeds int
        // from 2^30 rows)
        long long lo = 0, hi = (long long)rows + 1;
        while (lo < hi) {
          const long long mid = (lo + hi) >> 1;
          if (offsets[mid] < next) lo = mid + 1;
          else hi = mid;
        }
        if (lo > rows || offsets[lo] != next) f |= LOOPS_CSR_UNSORTED;
      }
    }
  }
  const unsigned long long any = __ballot(f != 0);
  if (any) {
    unsigned w = f;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) w |= __shfl_xor(w, m, 64);
    if ((threadIdx.x & 63) == 0) atomicOr(flags, w);
  }
}

hipError_t csr_structure_flags(const csr_args& A, hipStream_t s, unsigned* out) {
  *out = 0;
  if (A.rows == 0) {
    if (A.nnz != 0) *out = LOOPS_CSR_BAD_OFFSETS;
    return hipSuccess;
  }
  unsigned* d = nullptr;
  hipError_t e = hipMalloc(&d, sizeof(unsigned));
  if (e != hipSuccess) return e;
  e = hipMemsetAsync(d, 0, sizeof(unsigned), s);
  if (e == hipSuccess) {
    const long long n = A.rows > A.nnz ? A.rows : A.nnz;
    long long grid = (n + 255) / 256;
    if (grid > 16384) grid = 16384;
    csr_check_kernel<<<(unsigned)grid, 256, 0, s>>>(A.offsets, A.indices, A.rows, A.nnz,
                                                     A.cols, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, sizeof(unsigned), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d);
  return e;
}

}  // namespace loops_mi355x
