/**
 * @file timer.hxx
 * @brief Event timer (reference include/loops/util/timer.hxx:19-52):
 * start()/begin() record an event, stop()/end() record, wait and return
 * milliseconds.  Events are recorded on the given stream (default 0).
 */
#pragma once

#include <hip/hip_runtime_api.h>

namespace loops {
namespace util {

struct timer_t {
  float time = 0.f;

  explicit timer_t(hipStream_t stream = 0) : stream_(stream) {
    (void)hipEventCreate(&start_);
    (void)hipEventCreate(&stop_);
    (void)hipEventRecord(start_, stream_);
  }
  ~timer_t() {
    (void)hipEventDestroy(start_);
    (void)hipEventDestroy(stop_);
  }
  timer_t(const timer_t&) = delete;
  timer_t& operator=(const timer_t&) = delete;
  timer_t(timer_t&& o) noexcept
      : time(o.time), start_(o.start_), stop_(o.stop_), stream_(o.stream_) {
    o.start_ = nullptr;
    o.stop_ = nullptr;
  }

  void begin() { (void)hipEventRecord(start_, stream_); }
  void start() { begin(); }
  float end() {
    (void)hipEventRecord(stop_, stream_);
    (void)hipEventSynchronize(stop_);
    (void)hipEventElapsedTime(&time, start_, stop_);
    return milliseconds();
  }
  float stop() { return end(); }
  float seconds() const { return time * 1e-3f; }
  float milliseconds() const { return time; }

  /// Events for a library call to record around its kernel.
  hipEvent_t start_event() const { return start_; }
  hipEvent_t stop_event() const { return stop_; }

 private:
  hipEvent_t start_ = nullptr;
  hipEvent_t stop_ = nullptr;
  hipStream_t stream_ = 0;
};

}  // namespace util
}  // namespace loops
