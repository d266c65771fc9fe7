// C ABI of libloops_mi355x.so (declared in include/loops_capi.h): argument
// validation, defaults, workspace handling and dispatch to the kernels.
// Nothing here computes results on the host; there is no CPU fallback.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <mutex>
#include <new>

#include "loops_capi.h"
#include "loops_internal.hpp"

using namespace loops_mi355x;

struct loops_handle {
  int device = 0;
  int cu_count = 0;
  void* ws = nullptr;
  size_t ws_bytes = 0;
};

namespace {

thread_local int g_last_hip_error = 0;

loops_status_t hip_status(hipError_t e) {
  if (e == hipSuccess) return LOOPS_SUCCESS;
  g_last_hip_error = (int)e;
  if (e == hipErrorNotSupported) return LOOPS_ERROR_NOT_SUPPORTED;
  return LOOPS_ERROR_HIP;
}

// Default tunings (see the constants' use below).
constexpr int kMpDefaultTpb = 256;
constexpr int kMpDefaultIpt = 8;
constexpr int kRefBlock = 128;  // reference host entries use 128-thread blocks

loops_status_t check_csr(const loops_csr_view_t* A, csr_args* out) {
  if (!A) return LOOPS_ERROR_INVALID_VALUE;
  if (A->rows < 0 || A->cols < 0 || A->nnzs < 0) return LOOPS_ERROR_INVALID_VALUE;
  if (A->rows + A->nnzs > LOOPS_MAX_ROWS_PLUS_NNZS || A->cols >= (int64_t)INT_MAX)
    return LOOPS_ERROR_OVERFLOW;
  if (A->value_type != LOOPS_VALUE_F32 && A->value_type != LOOPS_VALUE_BF16 &&
      A->value_type != LOOPS_VALUE_F16)
    return LOOPS_ERROR_INVALID_VALUE;
  if (A->rows > 0 && !A->offsets) return LOOPS_ERROR_INVALID_VALUE;
  if (A->nnzs > 0 && (!A->indices || !A->values)) return LOOPS_ERROR_INVALID_VALUE;
  out->offsets = A->offsets;
  out->indices = A->indices;
  out->values = A->values;
  out->vtype = A->value_type;
  out->rows = (int)A->rows;
  out->cols = (int)A->cols;
  out->nnz = (int)A->nnzs;
  const uintptr_t vi = reinterpret_cast<uintptr_t>(A->indices);
  const uintptr_t vv = reinterpret_cast<uintptr_t>(A->values);
  const uintptr_t va = A->value_type == LOOPS_VALUE_F32 ? 16 : 8;
  out->aligned = (vi % 16 == 0) && (vv % va == 0);
  return LOOPS_SUCCESS;
}

struct resolved {
  int tpb = 0, ipt = 0, tpt = 0, grid = 0;
  long long units = 0;   // fix-up units (merge tiles / work_oriented threads)
  size_t ws_bytes = 0;
};

loops_status_t resolve_spmv(loops_handle_t h, loops_schedule_t s,
                            const csr_args& A, const loops_params_t* p,
                            resolved* r) {
  const loops_params_t zero{};
  if (!p) p = &zero;
  if (p->threads_per_block < 0 || p->items_per_thread < 0 ||
      p->threads_per_tile < 0 || p->grid_size < 0)
    return LOOPS_ERROR_INVALID_VALUE;
  const long long total = (long long)A.rows + A.nnz;
  switch (s) {
    case LOOPS_SCHEDULE_MERGE_PATH_FLAT: {
      r->tpb = p->threads_per_block ? p->threads_per_block : kMpDefaultTpb;
      r->ipt = p->items_per_thread ? p->items_per_thread : kMpDefaultIpt;
      if (!mp_spmv_supported(r->tpb, r->ipt)) return LOOPS_ERROR_NOT_SUPPORTED;
      const long long ti = (long long)r->tpb * r->ipt;
      r->units = (total + ti - 1) / ti;
      r->ws_bytes = r->units > 0 ? fixup_ws_bytes(r->units) : 0;
      return LOOPS_SUCCESS;
    }
    case LOOPS_SCHEDULE_WORK_ORIENTED: {
      r->tpb = p->threads_per_block ? p->threads_per_block : kRefBlock;
      if (!wo_spmv_supported(r->tpb)) return LOOPS_ERROR_NOT_SUPPORTED;
      r->grid = p->grid_size ? p->grid_size : 2 * (h ? h->cu_count : 256);
      r->units = (long long)r->grid * r->tpb;
      r->ws_bytes = total > 0 ? fixup_ws_bytes(r->units) : 0;
      return LOOPS_SUCCESS;
    }
    case LOOPS_SCHEDULE_GROUP_MAPPED: {
      r->tpb = p->threads_per_block ? p->threads_per_block : kRefBlock;
      r->tpt = p->threads_per_tile ? p->threads_per_tile : r->tpb;
      if (!gm_spmv_supported(r->tpb, r->tpt)) return LOOPS_ERROR_NOT_SUPPORTED;
      return LOOPS_SUCCESS;
    }
    case LOOPS_SCHEDULE_THREAD_MAPPED:
    case LOOPS_SCHEDULE_ORIGINAL: {
      r->tpb = p->threads_per_block ? p->threads_per_block : kRefBlock;
      if (r->tpb > 1024 || r->tpb % 64 != 0) return LOOPS_ERROR_NOT_SUPPORTED;
      return LOOPS_SUCCESS;
    }
    case LOOPS_SCHEDULE_BUCKETING:
      return LOOPS_ERROR_NOT_SUPPORTED;
  }
  return LOOPS_ERROR_INVALID_VALUE;
}

loops_status_t get_workspace(loops_handle_t h, size_t need, void* ws,
                             size_t ws_bytes, void** out) {
  *out = ws;
  if (need == 0) return LOOPS_SUCCESS;
  if (ws) return ws_bytes >= need ? LOOPS_SUCCESS : LOOPS_ERROR_WORKSPACE;
  if (!h) return LOOPS_ERROR_WORKSPACE;
  if (h->ws_bytes < need) {
    if (h->ws) {
      hipError_t e = hipFree(h->ws);
      h->ws = nullptr;
      h->ws_bytes = 0;
      if (e != hipSuccess) return hip_status(e);
    }
    size_t sz = need < (size_t(1) << 20) ? (size_t(1) << 20) : need;
    if (hipMalloc(&h->ws, sz) != hipSuccess) {
      h->ws = nullptr;
      return LOOPS_ERROR_ALLOCATION;
    }
    h->ws_bytes = sz;
  }
  *out = h->ws;
  return LOOPS_SUCCESS;
}

}  // namespace

namespace loops_mi355x {
// Scratch of loops_spmv_noalloc: the fix-up slots of merge_path_flat and
// work_oriented.  Each GPU holds its own copy until the process ends.  In
// host code the name is a host-side stand-in, not a device address: host
// code uses only the address scratch_address() looked up per device.
__device__ __align__(256) char scratch[LOOPS_NOALLOC_SCRATCH_BYTES];
}  // namespace loops_mi355x

namespace {

constexpr int kMaxScratchDevices = 64;
void* g_scratch_ptr[kMaxScratchDevices] = {};
std::mutex g_scratch_mu;

// Device address of `scratch` on the current device (looked up once per
// device, with that device current).
loops_status_t scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_status(e);
  if (dev < 0 || dev >= kMaxScratchDevices) return LOOPS_ERROR_NOT_SUPPORTED;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(loops_mi355x::scratch));
    if (e != hipSuccess) return hip_status(e);
    if (!p) return LOOPS_ERROR_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return LOOPS_SUCCESS;
}

}  // namespace

extern "C" {

loops_status_t loops_create(loops_handle_t* handle) {
  if (!handle) return LOOPS_ERROR_INVALID_VALUE;
  loops_handle* h = new (std::nothrow) loops_handle();
  if (!h) return LOOPS_ERROR_ALLOCATION;
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&h->cu_count, hipDeviceAttributeMultiprocessorCount,
                              h->device);
  if (e != hipSuccess) {
    delete h;
    return hip_status(e);
  }
  // set-up of loops_spmv_noalloc's scratch on this device
  void* sp = nullptr;
  const loops_status_t ss = scratch_address(&sp);
  if (ss != LOOPS_SUCCESS) {
    delete h;
    return ss;
  }
  *handle = h;
  return LOOPS_SUCCESS;
}

loops_status_t loops_destroy(loops_handle_t handle) {
  if (!handle) return LOOPS_ERROR_INVALID_VALUE;
  hipError_t e = hipSuccess;
  if (handle->ws) e = hipFree(handle->ws);
  delete handle;
  return hip_status(e);
}

const char* loops_status_string(loops_status_t status) {
  switch (status) {
    case LOOPS_SUCCESS: return "success";
    case LOOPS_ERROR_INVALID_VALUE: return "invalid value";
    case LOOPS_ERROR_OVERFLOW: return "index overflow (rows + nnzs must be at most 2^31 - 1 - 2^20)";
    case LOOPS_ERROR_WORKSPACE: return "workspace missing or too small";
    case LOOPS_ERROR_NOT_SUPPORTED: return "not supported";
    case LOOPS_ERROR_HIP: return "HIP runtime error";
    case LOOPS_ERROR_ALLOCATION: return "allocation failed";
  }
  return "unknown status";
}

const char* loops_schedule_name(loops_schedule_t schedule) {
  switch (schedule) {
    case LOOPS_SCHEDULE_MERGE_PATH_FLAT: return "merge_path_flat";
    case LOOPS_SCHEDULE_WORK_ORIENTED: return "work_oriented";
    case LOOPS_SCHEDULE_THREAD_MAPPED: return "thread_mapped";
    case LOOPS_SCHEDULE_GROUP_MAPPED: return "group_mapped";
    case LOOPS_SCHEDULE_BUCKETING: return "bucketing";
    case LOOPS_SCHEDULE_ORIGINAL: return "original";
  }
  return "unknown";
}

int loops_last_hip_error(void) { return g_last_hip_error; }

loops_status_t loops_device_cu_count(loops_handle_t handle, int32_t* cu_count) {
  if (!handle || !cu_count) return LOOPS_ERROR_INVALID_VALUE;
  *cu_count = handle->cu_count;
  return LOOPS_SUCCESS;
}

loops_status_t loops_spmv_workspace_size(loops_handle_t handle,
                                         loops_schedule_t schedule,
                                         const loops_csr_view_t* A,
                                         const loops_params_t* params,
                                         size_t* bytes) {
  if (!bytes) return LOOPS_ERROR_INVALID_VALUE;
  csr_args a;
  loops_status_t st = check_csr(A, &a);
  if (st != LOOPS_SUCCESS) return st;
  resolved r;
  st = resolve_spmv(handle, schedule, a, params, &r);
  if (st != LOOPS_SUCCESS) return st;
  *bytes = r.ws_bytes;
  return LOOPS_SUCCESS;
}

loops_status_t loops_spmv(loops_handle_t handle, loops_schedule_t schedule,
                          const loops_csr_view_t* A, const float* x, float* y,
                          const loops_params_t* params, void* workspace,
                          size_t workspace_bytes, hipStream_t stream) {
  csr_args a;
  loops_status_t st = check_csr(A, &a);
  if (st != LOOPS_SUCCESS) return st;
  if (a.rows > 0 && !y) return LOOPS_ERROR_INVALID_VALUE;
  if (a.nnz > 0 && !x) return LOOPS_ERROR_INVALID_VALUE;
  resolved r;
  st = resolve_spmv(handle, schedule, a, params, &r);
  if (st != LOOPS_SUCCESS) return st;
  void* ws = nullptr;
  st = get_workspace(handle, r.ws_bytes, workspace, workspace_bytes, &ws);
  if (st != LOOPS_SUCCESS) return st;
  hipEvent_t ev0 = params ? params->start_event : nullptr;
  hipEvent_t ev1 = params ? params->stop_event : nullptr;
  switch (schedule) {
    case LOOPS_SCHEDULE_MERGE_PATH_FLAT:
      return hip_status(launch_mp_spmv(a, x, y, r.tpb, r.ipt,
                                       params ? params->tile_coordinates : nullptr,
                                       carve_fixup_ws(ws, r.units), ev0, ev1,
                                       stream));
    case LOOPS_SCHEDULE_WORK_ORIENTED:
      return hip_status(launch_wo_spmv(a, x, y, r.tpb, r.grid,
                                       carve_fixup_ws(ws, r.units), ev0,
                                       ev1, stream));
    case LOOPS_SCHEDULE_GROUP_MAPPED:
      return hip_status(launch_gm_spmv(a, x, y, r.tpb, r.tpt, ev0, ev1, stream));
    case LOOPS_SCHEDULE_THREAD_MAPPED:
      return hip_status(launch_tm_spmv(a, x, y, r.tpb, false, ev0, ev1, stream));
    case LOOPS_SCHEDULE_ORIGINAL:
      return hip_status(launch_tm_spmv(a, x, y, r.tpb, true, ev0, ev1, stream));
    default:
      return LOOPS_ERROR_NOT_SUPPORTED;
  }
}

loops_status_t loops_spmv_noalloc(loops_handle_t handle, loops_schedule_t schedule,
                                  const loops_csr_view_t* A, const float* x, float* y,
                                  const loops_params_t* params, hipStream_t stream) {
  csr_args a;
  loops_status_t st = check_csr(A, &a);
  if (st != LOOPS_SUCCESS) return st;
  if (a.rows > 0 && !y) return LOOPS_ERROR_INVALID_VALUE;
  if (a.nnz > 0 && !x) return LOOPS_ERROR_INVALID_VALUE;
  resolved r;
  st = resolve_spmv(handle, schedule, a, params, &r);
  if (st != LOOPS_SUCCESS) return st;
  if (r.ws_bytes > (size_t)LOOPS_NOALLOC_SCRATCH_BYTES) return LOOPS_ERROR_WORKSPACE;
  void* ws = nullptr;
  if (r.ws_bytes > 0) {
    st = scratch_address(&ws);
    if (st != LOOPS_SUCCESS) return st;
  }
  // with a workspace of the size it needs, loops_spmv allocates nothing
  return loops_spmv(handle, schedule, A, x, y, params, ws, r.ws_bytes, stream);
}

}  // extern "C"

namespace {

// check_csr for loops_spmv_typed: value types F64 / I32 only.
loops_status_t check_typed_csr(const loops_csr_view_t* A, csr_args* out) {
  if (!A) return LOOPS_ERROR_INVALID_VALUE;
  if (A->value_type != LOOPS_VALUE_F64 && A->value_type != LOOPS_VALUE_I32)
    return LOOPS_ERROR_INVALID_VALUE;
  loops_csr_view_t v = *A;
  v.value_type = LOOPS_VALUE_F32;  // the shared checks of sizes and pointers
  const loops_status_t st = check_csr(&v, out);
  out->vtype = A->value_type;
  out->aligned = false;
  return st;
}

loops_status_t resolve_typed(loops_handle_t h, loops_schedule_t s, const csr_args& A,
                             const loops_params_t* p, typed_launch* L, long long* units) {
  const loops_params_t zero{};
  if (!p) p = &zero;
  if (p->threads_per_block < 0 || p->items_per_thread < 0 || p->threads_per_tile < 0 ||
      p->grid_size < 0)
    return LOOPS_ERROR_INVALID_VALUE;
  const bool wide = A.vtype == LOOPS_VALUE_F64;
  *L = typed_launch{};
  L->schedule = s;
  *units = 0;
  const long long total = (long long)A.rows + A.nnz;
  switch (s) {
    case LOOPS_SCHEDULE_MERGE_PATH_FLAT:
      L->tpb = p->threads_per_block ? p->threads_per_block : (wide ? 64 : 128);
      L->ipt = p->items_per_thread ? p->items_per_thread : (wide ? 3 : 5);
      *units = (total + (long long)L->tpb * L->ipt - 1) / ((long long)L->tpb * L->ipt);
      L->coords = p->tile_coordinates;
      break;
    case LOOPS_SCHEDULE_WORK_ORIENTED:
      L->tpb = p->threads_per_block ? p->threads_per_block : kRefBlock;
      L->grid = p->grid_size ? p->grid_size : 2 * (h ? h->cu_count : 256);
      *units = total > 0 ? (long long)L->grid * L->tpb : 0;
      break;
    case LOOPS_SCHEDULE_GROUP_MAPPED:
      L->tpb = p->threads_per_block ? p->threads_per_block : kRefBlock;
      L->tpt = p->threads_per_tile ? p->threads_per_tile : L->tpb;
      break;
    case LOOPS_SCHEDULE_THREAD_MAPPED:
    case LOOPS_SCHEDULE_ORIGINAL:
      L->tpb = p->threads_per_block ? p->threads_per_block : kRefBlock;
      break;
    default:
      return LOOPS_ERROR_NOT_SUPPORTED;
  }
  if (!typed_spmv_supported(s, L->tpb, L->ipt, L->tpt)) return LOOPS_ERROR_NOT_SUPPORTED;
  L->ev0 = p->start_event;
  L->ev1 = p->stop_event;
  return LOOPS_SUCCESS;
}

}  // namespace

extern "C" {

loops_status_t loops_spmv_typed_workspace_size(loops_handle_t handle,
                                               loops_schedule_t schedule,
                                               const loops_csr_view_t* A,
                                               const loops_params_t* params,
                                               size_t* bytes) {
  if (!bytes) return LOOPS_ERROR_INVALID_VALUE;
  csr_args a;
  loops_status_t st = check_typed_csr(A, &a);
  if (st != LOOPS_SUCCESS) return st;
  typed_launch L;
  long long units = 0;
  st = resolve_typed(handle, schedule, a, params, &L, &units);
  if (st != LOOPS_SUCCESS) return st;
  *bytes = units > 0 ? typed_fixup_bytes(units, a.vtype == LOOPS_VALUE_F64 ? 8 : 4) : 0;
  return LOOPS_SUCCESS;
}

loops_status_t loops_spmv_typed(loops_handle_t handle, loops_schedule_t schedule,
                                const loops_csr_view_t* A, const void* x, void* y,
                                const loops_params_t* params, void* workspace,
                                size_t workspace_bytes, hipStream_t stream) {
  csr_args a;
  loops_status_t st = check_typed_csr(A, &a);
  if (st != LOOPS_SUCCESS) return st;
  if (a.rows > 0 && !y) return LOOPS_ERROR_INVALID_VALUE;
  if (a.nnz > 0 && !x) return LOOPS_ERROR_INVALID_VALUE;
  typed_launch L;
  long long units = 0;
  st = resolve_typed(handle, schedule, a, params, &L, &units);
  if (st != LOOPS_SUCCESS) return st;
  const size_t need = units > 0 ? typed_fixup_bytes(units, a.vtype == LOOPS_VALUE_F64 ? 8 : 4) : 0;
  st = get_workspace(handle, need, workspace, workspace_bytes, &L.ws);
  if (st != LOOPS_SUCCESS) return st;
  return hip_status(launch_typed_spmv(L, a, x, y, stream));
}

loops_status_t loops_merge_path_num_tiles(int64_t rows, int64_t nnzs,
                                          int64_t items_per_tile,
                                          int64_t* num_tiles) {
  if (!num_tiles || rows < 0 || nnzs < 0 || items_per_tile <= 0)
    return LOOPS_ERROR_INVALID_VALUE;
  if (rows + nnzs > LOOPS_MAX_ROWS_PLUS_NNZS) return LOOPS_ERROR_OVERFLOW;
  *num_tiles = (rows + nnzs + items_per_tile - 1) / items_per_tile;
  return LOOPS_SUCCESS;
}

loops_status_t loops_merge_path_partition(const int32_t* offsets, int64_t rows,
                                          int64_t nnzs, int64_t items_per_tile,
                                          loops_coord_t* coords,
                                          hipStream_t stream) {
  int64_t tiles = 0;
  loops_status_t st = loops_merge_path_num_tiles(rows, nnzs, items_per_tile, &tiles);
  if (st != LOOPS_SUCCESS) return st;
  if (!coords || (rows > 0 && !offsets)) return LOOPS_ERROR_INVALID_VALUE;
  return hip_status(launch_mp_partition(offsets, (int)rows, (int)nnzs,
                                        items_per_tile, tiles + 1, coords, stream));
}

loops_status_t loops_work_oriented_partition(const int32_t* offsets,
                                             int64_t rows, int64_t nnzs,
                                             int64_t num_threads,
                                             loops_coord_t* coords,
                                             hipStream_t stream) {
  if (rows < 0 || nnzs < 0 || num_threads <= 0 || !coords ||
      (rows > 0 && !offsets))
    return LOOPS_ERROR_INVALID_VALUE;
  if (rows + nnzs > LOOPS_MAX_ROWS_PLUS_NNZS) return LOOPS_ERROR_OVERFLOW;
  return hip_status(launch_wo_partition(offsets, (int)rows, (int)nnzs,
                                        num_threads, coords, stream));
}

loops_status_t loops_csr_check(const loops_csr_view_t* A, uint32_t* flags,
                               hipStream_t stream) {
  if (!flags) return LOOPS_ERROR_INVALID_VALUE;
  csr_args a;
  const loops_status_t st = check_csr(A, &a);
  if (st != LOOPS_SUCCESS) return st;
  unsigned f = 0;
  const hipError_t e = csr_structure_flags(a, stream, &f);
  if (e != hipSuccess) return hip_status(e);
  *flags = f;
  return LOOPS_SUCCESS;
}

loops_status_t loops_spmv_plan_create(loops_handle_t handle,
                                      const loops_csr_view_t* A,
                                      const loops_plan_params_t* params,
                                      hipStream_t stream,
                                      loops_spmv_plan_t* plan) {
  (void)handle;
  if (!plan) return LOOPS_ERROR_INVALID_VALUE;
  csr_args a;
  loops_status_t st = check_csr(A, &a);
  if (st != LOOPS_SUCCESS) return st;
  const loops_plan_params_t zero{};
  const loops_plan_params_t* p = params ? params : &zero;
  if (p->panels < 0 || p->panels > kMaxPanels || p->layout < LOOPS_PLAN_LAYOUT_AUTO ||
      p->layout > LOOPS_PLAN_LAYOUT_BLOCKS || p->threads_per_block < 0 ||
      p->items_per_thread < 0)
    return LOOPS_ERROR_INVALID_VALUE;
  const bool panel_param = p->panels != 0 || p->threads_per_block != 0 || p->items_per_thread != 0;
  if (p->for_batches && (p->layout == LOOPS_PLAN_LAYOUT_PANELS || panel_param))
    return LOOPS_ERROR_INVALID_VALUE;
  // plans are validated up to this size (larger builds faulted the GPU)
  if ((long long)a.rows + a.nnz > LOOPS_MAX_PLAN_ROWS_PLUS_NNZS) return LOOPS_ERROR_NOT_SUPPORTED;
  // a plan bins entries by column and walks rows in column order: reject a
  // malformed CSR before building anything from it
  unsigned flags = 0;
  hipError_t ce = csr_structure_flags(a, stream, &flags);
  if (ce != hipSuccess) return hip_status(ce);
  if (flags) return LOOPS_ERROR_INVALID_VALUE;
  int layout = p->layout;
  // AUTO: row blocks from 2^12 columns and 2^19 nonzeros up, unless a
  // PANELS parameter is set.  Blocks against panels: 200k x 200k, 2M
  // nonzeros 20.2 -> 15.6 us; 10k x 1M, 5M 36.3 -> 26.2; 200k x 5k, 3M 19.0
  // -> 14.5; 30k x 30k, 600k 13.0 -> 11.3; narrow uniform matrices (40 or
  // 1,000 columns) and products under 2^19 nonzeros tie or favour panels.
  if (layout == LOOPS_PLAN_LAYOUT_AUTO)
    layout = p->for_batches || (a.cols >= (1 << 12) && a.nnz >= (1 << 19) && !panel_param)
                 ? LOOPS_PLAN_LAYOUT_BLOCKS
                 : LOOPS_PLAN_LAYOUT_PANELS;
  if (layout == LOOPS_PLAN_LAYOUT_BLOCKS) {
    loops_spmv_plan* pl = new (std::nothrow) loops_spmv_plan();
    if (!pl) return LOOPS_ERROR_ALLOCATION;
    pl->st.layout = LOOPS_PLAN_LAYOUT_BLOCKS;
    pl->st.rows = (int)A->rows;
    pl->rows = A->rows;
    pl->cols = A->cols;
    const hipError_t e = blk_plan_build(a, &pl->st.blk, p->for_batches != 0, stream);
    if (e != hipSuccess) {
      blk_plan_release(&pl->st.blk);
      delete pl;
      return e == hipErrorOutOfMemory ? LOOPS_ERROR_ALLOCATION : hip_status(e);
    }
    pl->st.P = pl->st.blk.S;
    pl->st.device_bytes = pl->st.blk.device_bytes;
    *plan = pl;
    return LOOPS_SUCCESS;
  }
  const int tpb = p->threads_per_block ? p->threads_per_block : 512;
  const int ipt = p->items_per_thread ? p->items_per_thread : 4;
  int panels = p->panels;
  if (panels == 0) {
    // one panel per 2 MiB of x (half an XCD's 4 MB L2), at most 8
    long long units = ((long long)a.cols * 4 + (2 << 20) - 1) >> 21;
    panels = units < 2 ? 1 : (units > kMaxPanels ? kMaxPanels : (int)units);
  }
  if (panels > 1 && !panels_supported(tpb, ipt)) return LOOPS_ERROR_NOT_SUPPORTED;
  if (panels == 1 && !mp_spmv_supported(tpb, ipt)) return LOOPS_ERROR_NOT_SUPPORTED;
  loops_spmv_plan* pl = new (std::nothrow) loops_spmv_plan();
  if (!pl) return LOOPS_ERROR_ALLOCATION;
  pl->st.aligned = a.aligned;
  pl->rows = A->rows;
  pl->cols = A->cols;
  hipError_t e = plan_build(a, panels, tpb, ipt, &pl->st, stream);
  if (e != hipSuccess) {
    plan_release(&pl->st);
    delete pl;
    return e == hipErrorOutOfMemory ? LOOPS_ERROR_ALLOCATION : hip_status(e);
  }
  *plan = pl;
  return LOOPS_SUCCESS;
}

loops_status_t loops_spmv_plan_execute(loops_spmv_plan_t plan, const float* x,
                                       float* y, hipEvent_t start_event,
                                       hipEvent_t stop_event,
                                       hipStream_t stream) {
  if (!plan) return LOOPS_ERROR_INVALID_VALUE;
  if ((plan->rows > 0 && !y) || (plan->cols > 0 && !x)) return LOOPS_ERROR_INVALID_VALUE;
  if (plan->st.layout == LOOPS_PLAN_LAYOUT_BLOCKS) {
    if (start_event) {
      const hipError_t e = hipEventRecord(start_event, stream);
      if (e != hipSuccess) return hip_status(e);
    }
    return hip_status(blk_plan_execute(plan->st.blk, x, y, stop_event, stream));
  }
  return hip_status(plan_execute(&plan->st, x, y, start_event, stop_event, stream));
}

loops_status_t loops_spmv_plan_info(loops_spmv_plan_t plan, int32_t* panels,
                                    int64_t* device_bytes) {
  if (!plan) return LOOPS_ERROR_INVALID_VALUE;
  if (panels) *panels = plan->st.P;
  if (device_bytes) *device_bytes = (int64_t)plan->st.device_bytes;
  return LOOPS_SUCCESS;
}

loops_status_t loops_spmv_plan_layout(loops_spmv_plan_t plan, int32_t* layout) {
  if (!plan || !layout) return LOOPS_ERROR_INVALID_VALUE;
  *layout = plan->st.layout;
  return LOOPS_SUCCESS;
}

loops_status_t loops_spmv_plan_blocks(loops_spmv_plan_t plan, int32_t* blocks,
                                      int32_t* compressed, int32_t* split_rows) {
  if (!plan) return LOOPS_ERROR_INVALID_VALUE;
  const bool blk = plan->st.layout == LOOPS_PLAN_LAYOUT_BLOCKS;
  if (blocks) *blocks = blk ? plan->st.blk.nb : 0;
  if (compressed) *compressed = blk && plan->st.blk.rlist != nullptr ? 1 : 0;
  if (split_rows) *split_rows = blk ? plan->st.blk.nsplit : 0;
  return LOOPS_SUCCESS;
}

loops_status_t loops_spmv_plan_destroy(loops_spmv_plan_t plan) {
  if (!plan) return LOOPS_ERROR_INVALID_VALUE;
  blk_plan_release(&plan->st.blk);
  plan_release(&plan->st);
  delete plan;
  return LOOPS_SUCCESS;
}

loops_status_t loops_batch_query(const loops_batch_item_t* items, int64_t n,
                                 loops_batch_info_t* info) {
  if (!info || n < 0 || (n > 0 && !items) || n >= INT_MAX) return LOOPS_ERROR_INVALID_VALUE;
  return batch_layout(items, n, info, nullptr, nullptr);
}

loops_status_t loops_batch_prepare(const loops_batch_item_t* items, int64_t n,
                                   void* device_workspace,
                                   size_t workspace_bytes, void* host_table,
                                   size_t table_bytes,
                                   loops_batch_info_t* info) {
  if (!info || !host_table || n < 0 || (n > 0 && !items) || n >= INT_MAX)
    return LOOPS_ERROR_INVALID_VALUE;
  loops_status_t st = batch_layout(items, n, info, nullptr, nullptr);
  if (st != LOOPS_SUCCESS) return st;
  if ((int64_t)table_bytes < info->table_bytes) return LOOPS_ERROR_INVALID_VALUE;
  if (!device_workspace || (int64_t)workspace_bytes < info->workspace_bytes)
    return LOOPS_ERROR_WORKSPACE;
  // Resident grid sizes of the current device, queried here (outside any
  // graph capture) and kept for the launches.
  hipError_t e = batch_warm_up();
  if (e != hipSuccess) return hip_status(e);
  return batch_layout(items, n, info, static_cast<batch_item_dev*>(host_table),
                      static_cast<char*>(device_workspace));
}

loops_status_t loops_batch_launch(const void* device_table,
                                  const loops_batch_info_t* info, int32_t grid,
                                  hipStream_t stream) {
  if (!info || (!device_table && info->work_items > 0) || grid < 0)
    return LOOPS_ERROR_INVALID_VALUE;
  return hip_status(batch_launch(device_table, info, grid, stream));
}

loops_status_t loops_spmm_workspace_size(loops_handle_t handle,
                                         loops_schedule_t schedule,
                                         const loops_csr_view_t* A, int64_t k,
                                         const loops_params_t* params,
                                         size_t* bytes) {
  (void)handle;
  if (!bytes || k < 0 || k > INT_MAX) return LOOPS_ERROR_INVALID_VALUE;
  csr_args a;
  loops_status_t st = check_csr(A, &a);
  if (st != LOOPS_SUCCESS) return st;
  const loops_params_t zero{};
  if (!params) params = &zero;
  switch (schedule) {
    case LOOPS_SCHEDULE_THREAD_MAPPED:
      *bytes = 0;
      return LOOPS_SUCCESS;
    case LOOPS_SCHEDULE_GROUP_MAPPED: {
      int tpb = params->threads_per_block ? params->threads_per_block : 256;
      if (!gm_spmm_supported(tpb)) return LOOPS_ERROR_NOT_SUPPORTED;
      *bytes = 0;
      return LOOPS_SUCCESS;
    }
    case LOOPS_SCHEDULE_MERGE_PATH_FLAT: {
      int tpb = params->threads_per_block ? params->threads_per_block : 256;
      int ipw = params->items_per_thread ? params->items_per_thread : 0;
      if (!mp_spmm_supported(tpb, ipw)) return LOOPS_ERROR_NOT_SUPPORTED;
      *bytes = mp_spmm_ws_need(a, (int)k, ipw);
      return LOOPS_SUCCESS;
    }
    default:
      return LOOPS_ERROR_NOT_SUPPORTED;
  }
}

loops_status_t loops_spmm(loops_handle_t handle, loops_schedule_t schedule,
                          const loops_csr_view_t* A, int64_t k, const float* B,
                          int64_t ldb, float* C, int64_t ldc,
                          const loops_params_t* params, void* workspace,
                          size_t workspace_bytes, hipStream_t stream) {
  csr_args a;
  loops_status_t st = check_csr(A, &a);
  if (st != LOOPS_SUCCESS) return st;
  if (k < 0 || k > INT_MAX || ldb < k || ldc < k) return LOOPS_ERROR_INVALID_VALUE;
  if (k > 0 && ((a.rows > 0 && !C) || (a.nnz > 0 && !B)))
    return LOOPS_ERROR_INVALID_VALUE;
  const loops_params_t zero{};
  const loops_params_t* p = params ? params : &zero;
  hipEvent_t ev0 = p->start_event, ev1 = p->stop_event;
  switch (schedule) {
    case LOOPS_SCHEDULE_THREAD_MAPPED: {
      int tpb = p->threads_per_block ? p->threads_per_block : kRefBlock;
      if (tpb > 1024 || tpb % 64 != 0) return LOOPS_ERROR_NOT_SUPPORTED;
      return hip_status(launch_tm_spmm(a, (int)k, B, ldb, C, ldc, tpb, ev0, ev1, stream));
    }
    case LOOPS_SCHEDULE_GROUP_MAPPED: {
      int tpb = p->threads_per_block ? p->threads_per_block : 1024;
      if (!gm_spmm_supported(tpb)) return LOOPS_ERROR_NOT_SUPPORTED;
      return hip_status(launch_gm_spmm(a, (int)k, B, ldb, C, ldc, tpb, ev0, ev1, stream));
    }
    case LOOPS_SCHEDULE_MERGE_PATH_FLAT: {
      int tpb = p->threads_per_block ? p->threads_per_block : 256;
      int ipw = p->items_per_thread ? p->items_per_thread : 0;
      if (!mp_spmm_supported(tpb, ipw) || !mp_spmm_fits((int)k, B, ldb, ipw))
        return LOOPS_ERROR_NOT_SUPPORTED;
      size_t need = mp_spmm_ws_need(a, (int)k, ipw);
      void* ws = nullptr;
      st = get_workspace(handle, need, workspace, workspace_bytes, &ws);
      if (st != LOOPS_SUCCESS) return st;
      return hip_status(launch_mp_spmm(a, (int)k, B, ldb, C, ldc, tpb, ipw, ws,
                                       ev0, ev1, stream));
    }
    default:
      return LOOPS_ERROR_NOT_SUPPORTED;
  }
}

}  // extern "C"
