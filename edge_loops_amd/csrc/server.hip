// Serving: continuous batching of SpMV / SpMM requests against CSR matrices
// that stay resident in HBM (loops_server_* in loops_capi.h).  No
// counterpart in the reference.
//
// Clients submit requests (matrix id, k right-hand sides in device memory);
// a server thread owns the GPU side:
//   admission  a request is admitted, in arrival order, once its output
//              (rows x k fp32) fits in the server's device arena, and its
//              share of a batch's table and workspace fits in a free slot;
//   batching   everything admissible when a slot (a stream with its own
//              table and workspace) is free goes into one batch: one
//              prologue, one persistent SpMV kernel, one persistent SpMM
//              kernel and the fix-ups cover all of its requests;
//   overlap    up to `slots` batches run at once on different streams, so
//              requests arriving while others are on the GPU are launched
//              without waiting for them;
//   completion an event per batch; the server thread stamps each request's
//              completion time and wakes waiting clients.  A result stays
//              in the arena until the client releases it.
// In plain mode the same loop launches every request on its own through
// loops_spmv / loops_spmm on one stream (the per-call baseline).  A request
// gives bit-identical results alone and in any mix: batch items are
// computed independently of each other.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "loops_capi.h"

namespace {

using steady = std::chrono::steady_clock;

inline size_t a256(size_t b) { return (b + 255) & ~size_t(255); }

// First-fit allocator of 256-byte granules over one device allocation.
class arena_t {
 public:
  void init(size_t bytes) {
    capacity_ = bytes;
    free_.clear();
    if (bytes) free_[0] = bytes;
  }
  bool alloc(size_t n, size_t* off) {
    n = n ? a256(n) : 256;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      *off = it->first;
      const size_t rest = it->second - n, o = it->first + n;
      free_.erase(it);
      if (rest) free_[o] = rest;
      used_ += n;
      if (used_ > high_) high_ = used_;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t n) {
    n = n ? a256(n) : 256;
    used_ -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t capacity() const { return capacity_; }
  size_t high_water() const { return high_; }

 private:
  std::map<size_t, size_t> free_;
  size_t capacity_ = 0, used_ = 0, high_ = 0;
};

enum req_state : int { RQ_QUEUED = 0, RQ_RUNNING = 1, RQ_DONE = 2, RQ_FAILED = 3 };

struct request_t {
  int64_t ticket = 0;
  loops_server_matrix_t m{};  // copied at submit
  int64_t k = 1;
  const float* x = nullptr;
  int64_t ldx = 1;
  hipEvent_t ready = nullptr;  // the producer of x (caller's event), or null
  loops_schedule_t schedule = LOOPS_SCHEDULE_MERGE_PATH_FLAT;
  int32_t ipt = 0;
  size_t y_off = 0, y_bytes = 0;
  float* y = nullptr;
  size_t ws_need = 0, table_need = 0;  // bounds of its share of a batch
  bool direct = false;  // BATCHED mode: launched alone ahead of its batch
  double t_submit = 0.0, t_done = 0.0;
  int state = RQ_QUEUED;
};

struct slot_t {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  void* host_table = nullptr;
  size_t host_table_bytes = 0;
  char* scratch = nullptr;  // device: batch table, then workspace
  std::vector<request_t*> reqs;
  bool busy = false;
};

// k -> index into loops_server_matrix_t.spmm_items_per_unit
int k_bucket(int64_t k) {
  int b = 0;
  int64_t lim = 2;
  while (b < 7 && k > lim) {
    lim *= 2;
    ++b;
  }
  return b;
}

void pick(const loops_server_matrix_t& m, int64_t k, loops_schedule_t* s, int32_t* ipt) {
  if (k == 1) {
    *s = (loops_schedule_t)m.spmv_schedule;
    *ipt = m.spmv_items_per_thread;
  } else {
    *s = LOOPS_SCHEDULE_MERGE_PATH_FLAT;
    *ipt = m.spmm_items_per_unit[k_bucket(k)];
  }
}

}  // namespace

struct loops_server {
  loops_server_config_t cfg{};
  int device = 0;
  char* arena_base = nullptr;
  arena_t arena;
  std::vector<loops_server_matrix_t> matrices;  // guarded by mu
  std::vector<slot_t> slots;
  loops_handle_t handle = nullptr;  // per-call launches
  void* plain_ws = nullptr;
  std::mutex mu;
  std::condition_variable cv_done;  // request completed
  std::condition_variable cv_work;  // request submitted / released
  std::deque<request_t*> pending;
  std::unordered_map<int64_t, request_t*> reqs;
  int64_t next_ticket = 1;
  std::atomic<bool> stop{false};
  std::thread th;
  loops_status_t fatal = LOOPS_SUCCESS;
  int fatal_hip = 0;
  steady::time_point t0;
  int64_t submitted = 0, completed = 0, batches = 0, batched_requests = 0;

  double now() const {
    return std::chrono::duration<double>(steady::now() - t0).count();
  }
  bool plain() const { return cfg.mode == LOOPS_SERVER_PLAIN; }
  bool reference_schedules() const {
    return (cfg.flags & LOOPS_SERVER_FLAG_REFERENCE_SCHEDULES) != 0;
  }
  bool batch_all() const { return (cfg.flags & LOOPS_SERVER_FLAG_BATCH_ALL) != 0; }
};

namespace {

// After an error the server accepts nothing more: queued requests fail too,
// so that no client waits for them.  Caller holds S->mu.
void fail_all_locked(loops_server* S, loops_status_t st, int hip) {
  if (S->fatal == LOOPS_SUCCESS) {
    S->fatal = st;
    S->fatal_hip = hip;
  }
  const double t = S->now();
  for (request_t* r : S->pending) {
    r->state = RQ_FAILED;
    r->t_done = t;
  }
  S->pending.clear();
}

// BATCHED mode: requests of at least this many work items ((rows + nnz) x
// k) are launched alone through loops_spmv / loops_spmm on the slot's
// stream, ahead of the batch of the others.  What a batch saves is launch
// overhead, small beside such a product (a 2M-nonzero k = 4 product runs
// ~47 us alone), while the batch adds its prologue (~8 us) and the
// persistent kernels' item search (~7 % on SpMM).  The result is the same
// either way (batch items are computed as alone).
constexpr int64_t kDirectWork = 1 << 20;

// The request's product through the per-call entry points, on `stream`
// with `ws` as workspace (PLAIN mode, and BATCHED mode's direct requests).
loops_status_t launch_one(loops_server* S, const request_t* r, void* ws, size_t ws_bytes,
                          hipStream_t stream) {
  loops_params_t p{};
  if (r->k == 1) {
    loops_schedule_t sched = r->schedule;
    if (S->reference_schedules()) {
      sched = LOOPS_SCHEDULE_MERGE_PATH_FLAT;  // library defaults
    } else if (sched == LOOPS_SCHEDULE_MERGE_PATH_FLAT) {
      p.threads_per_block = 256;
      p.items_per_thread = r->ipt ? r->ipt : 8;
    }
    return loops_spmv(S->handle, sched, &r->m.A, r->x, r->y, &p, ws, ws_bytes, stream);
  }
  if (S->reference_schedules())
    return loops_spmm(S->handle, LOOPS_SCHEDULE_THREAD_MAPPED, &r->m.A, r->k, r->x, r->ldx,
                      r->y, r->k, &p, ws, ws_bytes, stream);
  p.items_per_thread = r->ipt;
  return loops_spmm(S->handle, LOOPS_SCHEDULE_MERGE_PATH_FLAT, &r->m.A, r->k, r->x, r->ldx,
                    r->y, r->k, &p, ws, ws_bytes, stream);
}

// Workspace of launch_one for the request.
loops_status_t one_workspace(loops_server* S, const request_t* r, size_t* need) {
  loops_params_t p{};
  if (r->k == 1) {
    loops_schedule_t sched = r->schedule;
    if (S->reference_schedules()) {
      sched = LOOPS_SCHEDULE_MERGE_PATH_FLAT;
    } else if (sched == LOOPS_SCHEDULE_MERGE_PATH_FLAT) {
      p.threads_per_block = 256;
      p.items_per_thread = r->ipt ? r->ipt : 8;
    }
    return loops_spmv_workspace_size(nullptr, sched, &r->m.A, &p, need);
  }
  if (S->reference_schedules())
    return loops_spmm_workspace_size(nullptr, LOOPS_SCHEDULE_THREAD_MAPPED, &r->m.A, r->k, &p,
                                     need);
  p.items_per_thread = r->ipt;
  return loops_spmm_workspace_size(nullptr, LOOPS_SCHEDULE_MERGE_PATH_FLAT, &r->m.A, r->k, &p,
                                   need);
}

// Launches one batch (or, in plain mode, one request) on slot sl, after the
// producers of the requests' inputs (their ready events).
loops_status_t launch(loops_server* S, slot_t& sl, const std::vector<request_t*>& batch) {
  for (const request_t* r : batch)
    if (r->ready && hipStreamWaitEvent(sl.stream, r->ready, 0) != hipSuccess)
      return LOOPS_ERROR_HIP;
  if (S->plain()) {
    const loops_status_t st =
        launch_one(S, batch[0], S->plain_ws, (size_t)S->cfg.scratch_bytes, sl.stream);
    if (st != LOOPS_SUCCESS) return st;
    return hipEventRecord(sl.done, sl.stream) == hipSuccess ? LOOPS_SUCCESS
                                                            : LOOPS_ERROR_HIP;
  }
  // direct requests first, each with the slot's scratch as its workspace
  // (stream order separates their uses and the batch's)
  std::vector<const request_t*> rest;
  rest.reserve(batch.size());
  for (const request_t* r : batch) {
    if (!r->direct) {
      rest.push_back(r);
      continue;
    }
    const loops_status_t st =
        launch_one(S, r, sl.scratch, (size_t)S->cfg.scratch_bytes, sl.stream);
    if (st != LOOPS_SUCCESS) return st;
  }
  if (rest.empty())
    return hipEventRecord(sl.done, sl.stream) == hipSuccess ? LOOPS_SUCCESS
                                                            : LOOPS_ERROR_HIP;
  std::vector<loops_batch_item_t> items(rest.size());
  for (size_t i = 0; i < rest.size(); ++i) {
    const request_t* r = rest[i];
    loops_batch_item_t& it = items[i];
    it.A = r->m.A;
    it.k = r->k;
    it.x = r->x;
    it.ldx = r->k > 1 ? r->ldx : 1;
    it.y = r->y;
    it.ldy = r->k > 1 ? r->k : 1;
    it.schedule = r->schedule;
    it.items_per_thread = r->ipt;
  }
  loops_batch_info_t info{};
  loops_status_t st = loops_batch_query(items.data(), (int64_t)items.size(), &info);
  if (st != LOOPS_SUCCESS) return st;
  const size_t tb = a256((size_t)info.table_bytes);
  if (tb + (size_t)info.workspace_bytes > (size_t)S->cfg.scratch_bytes)
    return LOOPS_ERROR_WORKSPACE;  // excluded by the per-request bounds
  if (sl.host_table_bytes < (size_t)info.table_bytes) {
    // This slot's previous table copy has completed (the slot is free).
    if (sl.host_table) (void)hipHostFree(sl.host_table);
    sl.host_table = nullptr;
    size_t want = a256((size_t)info.table_bytes) * 2;
    if (hipHostMalloc(&sl.host_table, want, hipHostMallocDefault) != hipSuccess) {
      sl.host_table_bytes = 0;
      return LOOPS_ERROR_ALLOCATION;
    }
    sl.host_table_bytes = want;
  }
  st = loops_batch_prepare(items.data(), (int64_t)items.size(), sl.scratch + tb,
                           (size_t)S->cfg.scratch_bytes - tb, sl.host_table,
                           sl.host_table_bytes, &info);
  if (st != LOOPS_SUCCESS) return st;
  if (hipMemcpyAsync(sl.scratch, sl.host_table, (size_t)info.table_bytes,
                     hipMemcpyHostToDevice, sl.stream) != hipSuccess)
    return LOOPS_ERROR_HIP;
  st = loops_batch_launch(sl.scratch, &info, 0, sl.stream);
  if (st != LOOPS_SUCCESS) return st;
  return hipEventRecord(sl.done, sl.stream) == hipSuccess ? LOOPS_SUCCESS
                                                          : LOOPS_ERROR_HIP;
}

void server_loop(loops_server* S) {
  (void)hipSetDevice(S->device);
  const size_t cap = (size_t)S->cfg.max_batch_requests;
  std::vector<request_t*> batch;
  while (!S->stop.load(std::memory_order_relaxed)) {
    bool progressed = false, any_busy = false;
    // 1. retire finished batches
    for (slot_t& sl : S->slots) {
      if (!sl.busy) continue;
      const hipError_t q = hipEventQuery(sl.done);
      if (q == hipErrorNotReady) {
        any_busy = true;
        continue;
      }
      const double t = S->now();
      {
        std::lock_guard<std::mutex> lk(S->mu);
        for (request_t* r : sl.reqs) {
          r->state = q == hipSuccess ? RQ_DONE : RQ_FAILED;
          r->t_done = t;
        }
        S->completed += (int64_t)sl.reqs.size();
        if (q != hipSuccess) fail_all_locked(S, LOOPS_ERROR_HIP, (int)q);
      }
      sl.reqs.clear();
      sl.busy = false;
      progressed = true;
      S->cv_done.notify_all();
    }
    // 2. admit and launch into free slots
    for (slot_t& sl : S->slots) {
      if (sl.busy) continue;
      batch.clear();
      {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->fatal != LOOPS_SUCCESS) break;
        size_t share = 0;
        int direct = 0;
        while (!S->pending.empty() && batch.size() < cap) {
          request_t* r = S->pending.front();
          const size_t need = r->direct ? 0 : r->ws_need + r->table_need;
          if (!S->plain() && !batch.empty() && share + need > (size_t)S->cfg.scratch_bytes)
            break;
          // one request launched alone per slot: the next waits for a free
          // slot (and its stream) rather than behind this one
          if (r->direct && direct == 1) break;
          direct += r->direct ? 1 : 0;
          size_t off = 0;
          if (!S->arena.alloc(r->y_bytes, &off)) break;  // admission by memory
          r->y_off = off;
          r->y = reinterpret_cast<float*>(S->arena_base + off);
          r->state = RQ_RUNNING;
          share += need;
          batch.push_back(r);
          S->pending.pop_front();
          if (S->plain()) break;  // one request per launch
        }
      }
      if (batch.empty()) break;
      const loops_status_t st = launch(S, sl, batch);
      if (st != LOOPS_SUCCESS) {
        std::lock_guard<std::mutex> lk(S->mu);
        for (request_t* r : batch) {
          r->state = RQ_FAILED;
          r->t_done = S->now();
        }
        fail_all_locked(S, st, loops_last_hip_error());
        S->cv_done.notify_all();
        break;
      }
      sl.reqs = batch;
      sl.busy = true;
      any_busy = true;
      progressed = true;
      ++S->batches;
      S->batched_requests += (int64_t)batch.size();
    }
    if (progressed) continue;
    if (any_busy) {
      std::this_thread::yield();  // poll the in-flight batches
    } else {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait_for(lk, std::chrono::milliseconds(1), [&] {
        return S->stop.load() || (!S->pending.empty() && S->fatal == LOOPS_SUCCESS);
      });
    }
  }
}

void destroy_server(loops_server* S) {
  if (S->th.joinable()) {
    S->stop.store(true);
    S->cv_work.notify_all();
    S->th.join();
  }
  for (slot_t& sl : S->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
  }
  for (slot_t& sl : S->slots) {
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.host_table) (void)hipHostFree(sl.host_table);
    if (sl.scratch) (void)hipFree(sl.scratch);
    if (sl.stream && !(S->plain() && &sl != &S->slots[0])) (void)hipStreamDestroy(sl.stream);
  }
  if (S->plain_ws) (void)hipFree(S->plain_ws);
  if (S->arena_base) (void)hipFree(S->arena_base);
  if (S->handle) loops_destroy(S->handle);
  for (auto& kv : S->reqs) delete kv.second;
  delete S;
}

}  // namespace

extern "C" {

loops_status_t loops_server_create(const loops_server_config_t* config, loops_server_t* out) {
  if (!out) return LOOPS_ERROR_INVALID_VALUE;
  *out = nullptr;
  loops_server_config_t c = config ? *config : loops_server_config_t{};
  if (c.arena_bytes < 0 || c.scratch_bytes < 0 || c.slots < 0 || c.max_batch_requests < 0 ||
      (c.mode != LOOPS_SERVER_BATCHED && c.mode != LOOPS_SERVER_PLAIN) ||
      (c.flags & ~(LOOPS_SERVER_FLAG_REFERENCE_SCHEDULES | LOOPS_SERVER_FLAG_BATCH_ALL)) != 0 ||
      ((c.flags & LOOPS_SERVER_FLAG_REFERENCE_SCHEDULES) && c.mode != LOOPS_SERVER_PLAIN) ||
      ((c.flags & LOOPS_SERVER_FLAG_BATCH_ALL) && c.mode != LOOPS_SERVER_BATCHED))
    return LOOPS_ERROR_INVALID_VALUE;
  if (c.arena_bytes == 0) c.arena_bytes = (int64_t)1 << 30;
  if (c.scratch_bytes == 0) c.scratch_bytes = (int64_t)64 << 20;
  // batched: 3 slots (the served workload's capacity, 3,000 requests over
  // 64 matrices: 7,690 req/s with 2, 7,950-8,000 with 3 to 6)
  if (c.slots == 0) c.slots = c.mode == LOOPS_SERVER_PLAIN ? 64 : 3;
  if (c.max_batch_requests == 0) c.max_batch_requests = 512;
  if (c.slots > 1024) return LOOPS_ERROR_INVALID_VALUE;
  loops_server* S = new (std::nothrow) loops_server();
  if (!S) return LOOPS_ERROR_ALLOCATION;
  S->cfg = c;
  S->t0 = steady::now();
  auto fail = [&](loops_status_t st) {
    destroy_server(S);
    return st;
  };
  if (hipGetDevice(&S->device) != hipSuccess) return fail(LOOPS_ERROR_HIP);
  if (hipMalloc(reinterpret_cast<void**>(&S->arena_base), (size_t)c.arena_bytes) != hipSuccess)
    return fail(LOOPS_ERROR_ALLOCATION);
  S->arena.init((size_t)c.arena_bytes);
  S->slots.resize((size_t)c.slots);
  for (size_t i = 0; i < S->slots.size(); ++i) {
    slot_t& sl = S->slots[i];
    if (S->plain() && i > 0) {
      sl.stream = S->slots[0].stream;  // per-call launches share one stream
    } else if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess) {
      return fail(LOOPS_ERROR_HIP);
    }
    if (hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess)
      return fail(LOOPS_ERROR_HIP);
    if (!S->plain() &&
        hipMalloc(reinterpret_cast<void**>(&sl.scratch), (size_t)c.scratch_bytes) != hipSuccess)
      return fail(LOOPS_ERROR_ALLOCATION);
  }
  if (loops_create(&S->handle) != LOOPS_SUCCESS) return fail(LOOPS_ERROR_HIP);
  if (S->plain() && hipMalloc(&S->plain_ws, (size_t)c.scratch_bytes) != hipSuccess)
    return fail(LOOPS_ERROR_ALLOCATION);
  S->th = std::thread(server_loop, S);
  *out = S;
  return LOOPS_SUCCESS;
}

loops_status_t loops_server_destroy(loops_server_t S) {
  if (!S) return LOOPS_ERROR_INVALID_VALUE;
  destroy_server(S);
  return LOOPS_SUCCESS;
}

loops_status_t loops_server_add_matrix(loops_server_t S, const loops_server_matrix_t* m,
                                       int32_t* id) {
  if (!S || !m || !id) return LOOPS_ERROR_INVALID_VALUE;
  // Validate the matrix and its tunings once with a one-item batch query.
  loops_batch_item_t it{};
  it.A = m->A;
  it.k = 1;
  it.x = reinterpret_cast<const float*>(256);
  it.y = reinterpret_cast<float*>(256);
  it.ldx = it.ldy = 1;
  it.schedule = (loops_schedule_t)m->spmv_schedule;
  it.items_per_thread = m->spmv_items_per_thread;
  loops_batch_info_t info{};
  loops_status_t st = loops_batch_query(&it, 1, &info);
  if (st != LOOPS_SUCCESS) return st;
  for (int b = 0; b < 8; ++b) {
    if (m->spmm_items_per_unit[b] < 0) return LOOPS_ERROR_INVALID_VALUE;
  }
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->matrices.size() >= (size_t)INT32_MAX) return LOOPS_ERROR_OVERFLOW;
  *id = (int32_t)S->matrices.size();
  S->matrices.push_back(*m);
  return LOOPS_SUCCESS;
}

loops_status_t loops_server_submit(loops_server_t S, int32_t matrix, int64_t k, const float* x,
                                   int64_t ldx, hipEvent_t ready, int64_t* ticket) {
  if (!S || !ticket || k < 1) return LOOPS_ERROR_INVALID_VALUE;
  request_t* r = new (std::nothrow) request_t();
  if (!r) return LOOPS_ERROR_ALLOCATION;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->fatal != LOOPS_SUCCESS) {
      delete r;
      return S->fatal;
    }
    if (matrix < 0 || (size_t)matrix >= S->matrices.size()) {
      delete r;
      return LOOPS_ERROR_INVALID_VALUE;
    }
    r->m = S->matrices[(size_t)matrix];
  }
  r->k = k;
  r->x = x;
  r->ldx = k > 1 ? ldx : 1;
  r->ready = ready;
  pick(r->m, k, &r->schedule, &r->ipt);
  r->y_bytes = (size_t)r->m.A.rows * (size_t)k * sizeof(float);
  // Bounds of this request's share of a batch: the sizes of a batch of it alone.
  loops_batch_item_t it{};
  it.A = r->m.A;
  it.k = k;
  it.x = x;
  it.ldx = r->ldx;
  it.y = reinterpret_cast<float*>(256);  // 256-aligned, like arena outputs
  it.ldy = k > 1 ? k : 1;
  it.schedule = r->schedule;
  it.items_per_thread = r->ipt;
  loops_batch_info_t info{};
  loops_status_t st = loops_batch_query(&it, 1, &info);
  if (st == LOOPS_SUCCESS) {
    size_t need = 0;
    const loops_status_t so = one_workspace(S, r, &need);
    const bool fits = so == LOOPS_SUCCESS && need <= (size_t)S->cfg.scratch_bytes;
    if (S->plain()) {
      st = so != LOOPS_SUCCESS ? so : fits ? LOOPS_SUCCESS : LOOPS_ERROR_WORKSPACE;
    } else {
      const int64_t work = ((int64_t)r->m.A.rows + r->m.A.nnzs) * k;
      r->direct = !S->batch_all() && fits && work >= kDirectWork;
    }
  }
  if (st != LOOPS_SUCCESS) {
    delete r;
    return st;
  }
  r->ws_need = (size_t)info.workspace_bytes;
  r->table_need = a256((size_t)info.table_bytes);
  if (r->y_bytes > S->arena.capacity() ||
      (!S->plain() && r->ws_need + r->table_need > (size_t)S->cfg.scratch_bytes)) {
    delete r;
    return LOOPS_ERROR_WORKSPACE;  // could never be admitted
  }
  {
    std::lock_guard<std::mutex> lk(S->mu);
    r->ticket = S->next_ticket++;
    r->t_submit = S->now();
    S->reqs[r->ticket] = r;
    S->pending.push_back(r);
    ++S->submitted;
    *ticket = r->ticket;
  }
  S->cv_work.notify_one();
  return LOOPS_SUCCESS;
}

loops_status_t loops_server_test(loops_server_t S, int64_t ticket, int32_t* done, float** y,
                                 double* submit_s, double* done_s) {
  if (!S || !done) return LOOPS_ERROR_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(ticket);
  if (it == S->reqs.end()) return LOOPS_ERROR_INVALID_VALUE;
  const request_t* r = it->second;
  if (r->state == RQ_FAILED) return S->fatal != LOOPS_SUCCESS ? S->fatal : LOOPS_ERROR_HIP;
  *done = r->state == RQ_DONE ? 1 : 0;
  if (y) *y = r->state == RQ_DONE ? r->y : nullptr;
  if (submit_s) *submit_s = r->t_submit;
  if (done_s) *done_s = r->t_done;
  return LOOPS_SUCCESS;
}

loops_status_t loops_server_wait(loops_server_t S, int64_t ticket, float** y, double* submit_s,
                                 double* done_s) {
  if (!S) return LOOPS_ERROR_INVALID_VALUE;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(ticket);
  if (it == S->reqs.end()) return LOOPS_ERROR_INVALID_VALUE;
  request_t* r = it->second;
  S->cv_done.wait(lk, [&] { return r->state >= RQ_DONE; });
  if (r->state == RQ_FAILED) return S->fatal != LOOPS_SUCCESS ? S->fatal : LOOPS_ERROR_HIP;
  if (y) *y = r->y;
  if (submit_s) *submit_s = r->t_submit;
  if (done_s) *done_s = r->t_done;
  return LOOPS_SUCCESS;
}

loops_status_t loops_server_release(loops_server_t S, int64_t ticket) {
  if (!S) return LOOPS_ERROR_INVALID_VALUE;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return LOOPS_ERROR_INVALID_VALUE;
    request_t* r = it->second;
    if (r->state < RQ_DONE) return LOOPS_ERROR_INVALID_VALUE;  // still queued or running
    if (r->y) S->arena.release(r->y_off, r->y_bytes);
    S->reqs.erase(it);
    delete r;
  }
  S->cv_work.notify_one();  // memory freed: admission may proceed
  return LOOPS_SUCCESS;
}

loops_status_t loops_server_now(loops_server_t S, double* now_s) {
  if (!S || !now_s) return LOOPS_ERROR_INVALID_VALUE;
  *now_s = S->now();
  return LOOPS_SUCCESS;
}

loops_status_t loops_server_stats(loops_server_t S, loops_server_stats_t* st) {
  if (!S || !st) return LOOPS_ERROR_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(S->mu);
  st->submitted = S->submitted;
  st->completed = S->completed;
  st->batches = S->batches;
  st->batched_requests = S->batched_requests;
  st->arena_bytes = (int64_t)S->arena.capacity();
  st->arena_high_water = (int64_t)S->arena.high_water();
  return LOOPS_SUCCESS;
}

loops_status_t loops_server_replay(loops_server_t S, const loops_server_request_t* reqs,
                                   int64_t n, const double* arrival_s, double* latency_s,
                                   double* makespan_s) {
  if (!S || n < 0 || (n > 0 && (!reqs || !arrival_s || !latency_s))) return LOOPS_ERROR_INVALID_VALUE;
  for (int64_t i = 1; i < n; ++i)
    if (arrival_s[i] < arrival_s[i - 1]) return LOOPS_ERROR_INVALID_VALUE;
  std::vector<std::pair<int64_t, int64_t>> open;  // (request index, ticket)
  open.reserve((size_t)n);
  const double start = S->now();
  double last_done = start;
  int64_t next = 0, finished = 0;
  loops_status_t err = LOOPS_SUCCESS;
  while (finished < n) {
    const double t = S->now() - start;
    while (err == LOOPS_SUCCESS && next < n && arrival_s[next] <= t) {
      int64_t tk = 0;
      const loops_server_request_t& q = reqs[next];
      err = loops_server_submit(S, q.matrix, q.k, q.x, q.ldx, nullptr, &tk);
      if (err != LOOPS_SUCCESS) break;
      open.emplace_back(next, tk);
      ++next;
    }
    if (err != LOOPS_SUCCESS) n = next;  // drain what was submitted, then report
    size_t w = 0;
    for (size_t i = 0; i < open.size(); ++i) {
      int32_t done = 0;
      double ds = 0.0;
      loops_status_t st = loops_server_test(S, open[i].second, &done, nullptr, nullptr, &ds);
      if (st != LOOPS_SUCCESS) return st;
      if (done) {
        latency_s[open[i].first] = ds - (start + arrival_s[open[i].first]);
        if (ds > last_done) last_done = ds;
        loops_server_release(S, open[i].second);
        ++finished;
      } else {
        open[w++] = open[i];
      }
    }
    open.resize(w);
    if (finished < n && (next >= n || arrival_s[next] > S->now() - start))
      std::this_thread::yield();
  }
  if (makespan_s) *makespan_s = last_done - start - (n > 0 ? arrival_s[0] : 0.0);
  return err;
}

}  // extern "C"
