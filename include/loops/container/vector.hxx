/**
 * @file vector.hxx
 * @brief Owned 1-D array in host or device memory (reference
 * include/loops/container/vector.hxx:31-42, which aliases thrust vectors).
 *
 * vector_t<T, device> owns a hipMalloc'd buffer (value-initialised to zero
 * like a thrust::device_vector); vector_t<T, host> owns host memory.
 * Copy construction between spaces copies the contents (hipMemcpy).
 * data() returns a pointer wrapper with .get() for device vectors and a raw
 * pointer for host vectors, as thrust does.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstring>
#include <initializer_list>
#include <utility>
#include <vector>

#include <loops/error.hxx>
#include <loops/memory.hxx>

namespace loops {

template <typename type_t, memory_space_t space = memory_space_t::device>
class vector_t;

template <typename type_t>
class vector_t<type_t, memory_space_t::host> {
 public:
  using value_type = type_t;
  vector_t() = default;
  explicit vector_t(std::size_t n) : v_(n, type_t()) {}
  vector_t(std::size_t n, type_t value) : v_(n, value) {}
  vector_t(std::initializer_list<type_t> l) : v_(l) {}
  template <memory_space_t other>
  vector_t(const vector_t<type_t, other>& rhs) : v_(rhs.size()) {
    if (rhs.size())
      error::throw_if_exception(
          hipMemcpy(v_.data(), rhs.raw(), rhs.size() * sizeof(type_t),
                    hipMemcpyDefault),
          "vector_t copy to host");
  }
  vector_t(const vector_t&) = default;
  vector_t(vector_t&&) = default;
  vector_t& operator=(const vector_t&) = default;
  vector_t& operator=(vector_t&&) = default;

  std::size_t size() const { return v_.size(); }
  void resize(std::size_t n) { v_.resize(n); }
  type_t* data() { return v_.data(); }
  const type_t* data() const { return v_.data(); }
  const type_t* raw() const { return v_.data(); }
  type_t* raw() { return v_.data(); }
  type_t& operator[](std::size_t i) { return v_[i]; }
  const type_t& operator[](std::size_t i) const { return v_[i]; }
  type_t* begin() { return v_.data(); }
  type_t* end() { return v_.data() + v_.size(); }

 private:
  std::vector<type_t> v_;
};

template <typename type_t>
class vector_t<type_t, memory_space_t::device> {
 public:
  using value_type = type_t;
  vector_t() = default;
  explicit vector_t(std::size_t n) { allocate(n, true); }
  template <memory_space_t other>
  vector_t(const vector_t<type_t, other>& rhs) {
    allocate(rhs.size(), false);
    if (n_)
      error::throw_if_exception(
          hipMemcpy(p_, rhs.raw(), n_ * sizeof(type_t), hipMemcpyDefault),
          "vector_t copy to device");
  }
  vector_t(const vector_t& rhs) : vector_t() {
    allocate(rhs.size(), false);
    if (n_)
      error::throw_if_exception(
          hipMemcpy(p_, rhs.p_, n_ * sizeof(type_t), hipMemcpyDeviceToDevice),
          "vector_t copy");
  }
  vector_t(vector_t&& rhs) noexcept : p_(rhs.p_), n_(rhs.n_) {
    rhs.p_ = nullptr;
    rhs.n_ = 0;
  }
  vector_t& operator=(vector_t&& rhs) noexcept {
    if (this != &rhs) {
      release();
      p_ = rhs.p_;
      n_ = rhs.n_;
      rhs.p_ = nullptr;
      rhs.n_ = 0;
    }
    return *this;
  }
  vector_t& operator=(const vector_t& rhs) {
    if (this != &rhs) {
      vector_t tmp(rhs);
      *this = std::move(tmp);
    }
    return *this;
  }
  ~vector_t() { release(); }

  std::size_t size() const { return n_; }
  void resize(std::size_t n) {
    vector_t tmp(n);
    std::size_t keep = n < n_ ? n : n_;
    if (keep)
      error::throw_if_exception(
          hipMemcpy(tmp.p_, p_, keep * sizeof(type_t), hipMemcpyDeviceToDevice),
          "vector_t resize");
    *this = std::move(tmp);
  }
  memory::device_ptr<type_t> data() const { return {p_}; }
  type_t* raw() const { return p_; }
  type_t* begin() const { return p_; }
  type_t* end() const { return p_ + n_; }

 private:
  void allocate(std::size_t n, bool zero) {
    n_ = n;
    if (n) {
      error::throw_if_exception(hipMalloc(&p_, n * sizeof(type_t)),
                                "vector_t allocation");
      if (zero)
        error::throw_if_exception(hipMemset(p_, 0, n * sizeof(type_t)),
                                  "vector_t zero fill");
    }
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  type_t* p_ = nullptr;
  std::size_t n_ = 0;
};

template <typename type_t>
using host_vector_t = vector_t<type_t, memory_space_t::host>;
template <typename type_t>
using device_vector_t = vector_t<type_t, memory_space_t::device>;

}  // namespace loops
