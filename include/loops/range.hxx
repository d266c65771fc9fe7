/**
 * @file range.hxx
 * @brief Range-for proxies returned by the schedules' tiles()/atoms().
 *
 * Same interface as the reference's include/loops/range.hxx:53-216:
 * range(b, e), range(b) (unbounded), .step(s), indices(container | array |
 * initializer list).
 * Termination of a stepped range is "current passed end" (range.hxx:78-80);
 * the unit-step range here also stops once current >= end, so a range whose
 * begin exceeds its end is empty instead of running forever.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <initializer_list>
#include <type_traits>

namespace loops {

template <typename type_t>
struct range_proxy {
  struct iter {
    type_t current;
    __host__ __device__ iter(type_t c) : current(c) {}
    __host__ __device__ type_t operator*() const { return current; }
    __host__ __device__ iter& operator++() {
      ++current;
      return *this;
    }
    __host__ __device__ bool operator==(iter const& o) const {
      return current >= o.current;
    }
    __host__ __device__ bool operator!=(iter const& o) const {
      return !(*this == o);
    }
  };

  struct step_range_proxy {
    struct iter {
      type_t current;
      type_t stride;
      __host__ __device__ iter(type_t c, type_t s) : current(c), stride(s) {}
      __host__ __device__ type_t operator*() const { return current; }
      __host__ __device__ iter& operator++() {
        current += stride;
        return *this;
      }
      __host__ __device__ bool operator==(iter const& o) const {
        return stride > 0 ? current >= o.current : current < o.current;
      }
      __host__ __device__ bool operator!=(iter const& o) const {
        return !(*this == o);
      }
    };
    __host__ __device__ step_range_proxy(type_t b, type_t e, type_t s)
        : begin_(b, s), end_(e, s) {}
    __host__ __device__ iter begin() const { return begin_; }
    __host__ __device__ iter end() const { return end_; }

   private:
    iter begin_;
    iter end_;
  };

  __host__ __device__ range_proxy(type_t b, type_t e) : begin_(b), end_(e) {}
  __host__ __device__ step_range_proxy step(type_t s) const {
    return {begin_.current, end_.current, s};
  }
  __host__ __device__ iter begin() const { return begin_; }
  __host__ __device__ iter end() const { return end_; }

 private:
  iter begin_;
  iter end_;
};

template <typename type_t>
struct infinite_range_proxy {
  struct iter {
    type_t current;
    type_t stride;
    __host__ __device__ iter(type_t c = type_t(), type_t s = type_t(1))
        : current(c), stride(s) {}
    __host__ __device__ type_t operator*() const { return current; }
    __host__ __device__ iter& operator++() {
      current += stride;
      return *this;
    }
    __host__ __device__ bool operator==(iter const&) const { return false; }
    __host__ __device__ bool operator!=(iter const&) const { return true; }
  };
  struct step_range_proxy {
    __host__ __device__ step_range_proxy(type_t b, type_t s) : begin_(b, s) {}
    __host__ __device__ iter begin() const { return begin_; }
    __host__ __device__ iter end() const { return iter(); }

   private:
    iter begin_;
  };
  __host__ __device__ infinite_range_proxy(type_t b) : begin_(b) {}
  __host__ __device__ step_range_proxy step(type_t s) const {
    return step_range_proxy(begin_.current, s);
  }
  __host__ __device__ iter begin() const { return begin_; }
  __host__ __device__ iter end() const { return iter(); }

 private:
  iter begin_;
};

template <typename type_t>
__host__ __device__ range_proxy<type_t> range(type_t begin, type_t end) {
  return {begin, end};
}

template <typename type_t>
__host__ __device__ infinite_range_proxy<type_t> range(type_t begin) {
  return {begin};
}

template <typename C>
__host__ __device__ auto indices(C const& cont)
    -> range_proxy<decltype(cont.size())> {
  return {0, cont.size()};
}

template <typename type_t, std::size_t N>
__host__ __device__ range_proxy<std::size_t> indices(type_t (&)[N]) {
  return {0, N};
}

/// indices({a, b, c}): 0 .. list size (range.hxx:220-224).
template <typename type_t>
__host__ __device__ range_proxy<typename std::initializer_list<type_t>::size_type>
indices(std::initializer_list<type_t>&& cont) {
  return {0, cont.size()};
}

}  // namespace loops
